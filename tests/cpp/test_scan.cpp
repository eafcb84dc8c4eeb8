// Compiled tests of the C++ host interface (include/morphogen_hip.hpp), shaped
// like the reference's own kernel tests (crates/morphogen-gpu-dpf/src/
// kernel.rs:1390-1402 make_test_pages, :1443-1564 recovery / non-power-of-two
// / tile boundary / timing, :1566-1690 GPU fused recovery and tiled batch ==
// default batch). Every GPU answer is also compared byte for byte with the
// oracle's tiled CPU scan (oracle/oracle.h, test infrastructure).
//
//   test_scan_cpp --cpu   checks that need no GPU (parameters, key parsing,
//                         keygen == oracle keygen, no-device error)
//   test_scan_cpp --gpu   the GPU tests
// Exit status 0 when every check passes; each failure prints its case.
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

#include "../../include/morphogen_hip.hpp"
#include "../../oracle/oracle.h"

using namespace morphogen;

namespace {

int g_failures = 0;

#define CHECK(cond, ...)                                             \
  do {                                                               \
    if (!(cond)) {                                                   \
      std::fprintf(stderr, "FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
      std::fprintf(stderr, __VA_ARGS__);                             \
      std::fprintf(stderr, "\n");                                    \
      g_failures++;                                                  \
    }                                                                \
  } while (0)

template <class E>
bool throws(const std::function<void()>& f) {
  try {
    f();
  } catch (const E&) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

// make_test_pages (kernel.rs:1390-1402): page i starts with i (LE, 2 bytes),
// then byte j = (i + j) & 0xFF.
std::vector<uint8_t> make_test_pages(size_t num_pages) {
  std::vector<uint8_t> flat(num_pages * PAGE_SIZE_BYTES);
  for (size_t i = 0; i < num_pages; i++) {
    uint8_t* p = flat.data() + i * PAGE_SIZE_BYTES;
    p[0] = (uint8_t)(i & 0xFF);
    p[1] = (uint8_t)((i >> 8) & 0xFF);
    for (size_t j = 2; j < PAGE_SIZE_BYTES; j++) p[j] = (uint8_t)((i + j) & 0xFF);
  }
  return flat;
}

std::vector<uint8_t> page_of(const std::vector<uint8_t>& flat, size_t i) {
  return std::vector<uint8_t>(flat.begin() + i * PAGE_SIZE_BYTES,
                              flat.begin() + (i + 1) * PAGE_SIZE_BYTES);
}

std::vector<uint8_t> xor_of(const std::vector<uint8_t>& a, const std::vector<uint8_t>& b) {
  std::vector<uint8_t> r(a.size());
  for (size_t i = 0; i < a.size(); i++) r[i] = a[i] ^ b[i];
  return r;
}

// The oracle's answer of one wire key over the pages.
std::vector<uint8_t> oracle_answer(const ChaChaKey& k, const std::vector<uint8_t>& flat) {
  orc_key ok;
  const auto& w = k.to_bytes();
  if (orc_key_from_bytes(w.data(), w.size(), &ok) != ORC_OK) return {};
  std::vector<uint8_t> out(PAGE_SIZE_BYTES);
  orc_scan_tiled(&ok, 1, flat.data(), flat.size() / PAGE_SIZE_BYTES, PAGE_SIZE_BYTES, 8,
                 out.data());
  return out;
}

using Query = std::array<ChaChaKey, 3>;

struct Pair {
  Query a, b;  // party 0 and party 1 keys of three targets
};

Pair key_pairs(uint32_t d, const std::array<uint64_t, 3>& targets, uint64_t seed) {
  const ChaChaParams p = ChaChaParams::create(d);
  auto k0 = generate_chacha_dpf_keys(p, targets[0], seed);
  auto k1 = generate_chacha_dpf_keys(p, targets[1], seed + 1);
  auto k2 = generate_chacha_dpf_keys(p, targets[2], seed + 2);
  return {{k0.first, k1.first, k2.first}, {k0.second, k1.second, k2.second}};
}

void check_recovery(const char* name, const GpuScanner& s, uint32_t d, size_t num_pages,
                    const std::array<uint64_t, 3>& targets) {
  std::printf("case %s: d=%u pages=%zu\n", name, d, num_pages);
  std::fflush(stdout);
  const auto flat = make_test_pages(num_pages);
  GpuPageMatrix db = GpuPageMatrix::create(s, flat.data(), flat.size());
  const Pair kp = key_pairs(d, targets, 100 + d);
  const PirResult r0 = s.scan(db, {&kp.a[0], &kp.a[1], &kp.a[2]});
  const PirResult r1 = s.scan(db, {&kp.b[0], &kp.b[1], &kp.b[2]});
  const std::vector<uint8_t>* p0[3] = {&r0.page0, &r0.page1, &r0.page2};
  const std::vector<uint8_t>* p1[3] = {&r1.page0, &r1.page1, &r1.page2};
  for (int i = 0; i < 3; i++) {
    CHECK(xor_of(*p0[i], *p1[i]) == page_of(flat, targets[i]), "%s: target %d", name, i);
    CHECK(*p0[i] == oracle_answer(kp.a[i], flat), "%s: server 0 key %d != oracle", name, i);
    CHECK(*p1[i] == oracle_answer(kp.b[i], flat), "%s: server 1 key %d != oracle", name, i);
  }
  CHECK(r0.verif0 == std::vector<uint8_t>(16) && r0.verif2 == std::vector<uint8_t>(16),
        "%s: verifier words", name);
}

void gpu_tests() {
  GpuScanner scanner(0);
  // gpu_fused_3dpf_recovers_all_targets (kernel.rs:1566-1611).
  check_recovery("gpu_fused_3dpf_recovers_all_targets", scanner, 8, 256, {10, 100, 200});
  // fused_handles_non_power_of_two_pages (kernel.rs:1510-1534).
  check_recovery("fused_handles_non_power_of_two_pages", scanner, 10, 500, {5, 150, 300});
  // tile_processing_correctness (kernel.rs:1536-1564).
  check_recovery("tile_processing_correctness", scanner, 12, 4096, {2048, 100, 3000});

  // gpu_tiled_batch_matches_default_optimized_outputs (kernel.rs:1613-1689),
  // plus the per-query launches, a graph replay and the oracle.
  std::printf("case gpu_tiled_batch_matches_default_optimized_outputs\n");
  std::fflush(stdout);
  const size_t num_pages = 256;
  const auto flat = make_test_pages(num_pages);
  GpuPageMatrix db = GpuPageMatrix::create(scanner, flat.data(), flat.size());
  const ChaChaParams p = ChaChaParams::create(8);
  std::vector<Query> queries;
  for (size_t i = 0; i < 32; i++) {
    const uint64_t t0 = i % num_pages, t1 = (i * 7 + 13) % num_pages, t2 = (i * 11 + 29) % num_pages;
    queries.push_back({generate_chacha_dpf_keys(p, t0, 1000 + 3 * i).first,
                       generate_chacha_dpf_keys(p, t1, 1001 + 3 * i).first,
                       generate_chacha_dpf_keys(p, t2, 1002 + 3 * i).first});
  }
  for (size_t q : {1, 2, 4, 8, 16, 32}) {
    std::vector<Query> slice(queries.begin(), queries.begin() + q);
    const auto base = scanner.scan_batch_optimized(db, slice);
    const auto tiled = scanner.scan_batch_optimized_tiled(db, slice, 4);
    const auto graph = scanner.scan_batch_optimized_with_graph(db, slice, true);
    const auto multi = scanner.scan_batch_single_query_multistream_optimized(db, slice, 4);
    CHECK(base.size() == q && tiled.size() == q && graph.size() == q && multi.size() == q,
          "q=%zu: result count", q);
    for (size_t i = 0; i < q && i < base.size(); i++) {
      CHECK(base[i].page0 == tiled[i].page0 && base[i].page1 == tiled[i].page1 &&
                base[i].page2 == tiled[i].page2 && base[i].verif0 == tiled[i].verif0,
            "q=%zu idx=%zu: tiled != default", q, i);
      CHECK(base[i].page0 == graph[i].page0 && base[i].page2 == graph[i].page2,
            "q=%zu idx=%zu: graph != eager", q, i);
      CHECK(base[i].page0 == multi[i].page0 && base[i].page1 == multi[i].page1 &&
                base[i].page2 == multi[i].page2,
            "q=%zu idx=%zu: per-query != batch", q, i);
    }
  }
  const auto all = scanner.scan_batch(db, queries);
  for (size_t i = 0; i < queries.size(); i += 7) {
    CHECK(all[i].page0 == oracle_answer(queries[i][0], flat), "batch idx=%zu != oracle", i);
    CHECK(all[i].page2 == oracle_answer(queries[i][2], flat), "batch idx=%zu != oracle", i);
  }
  // fused_timing_is_recorded (kernel.rs:1495-1508).
  CHECK(all[0].timing.total_ns > 0 && all[0].timing.kernel_ns > 0, "timing recorded");

  // ChaChaKey::{eval, eval_subtree, full_eval} on the device == the oracle.
  std::printf("case key evaluation\n");
  std::fflush(stdout);
  {
    const auto kp = generate_chacha_dpf_keys(ChaChaParams::create(12), 1234, 55);
    orc_key ok;
    const auto& w = kp.first.to_bytes();
    CHECK(orc_key_from_bytes(w.data(), w.size(), &ok) == ORC_OK, "oracle key");
    const auto full = scanner.full_eval(kp.first);
    const auto full1 = scanner.full_eval(kp.second);
    std::vector<uint32_t> ref(4u << 12);
    CHECK(orc_full_eval(&ok, ref.data()) == ORC_OK, "oracle full_eval");
    CHECK(full.size() == 4096 && std::memcmp(full.data(), ref.data(), 4096 * 16) == 0,
          "full_eval != oracle");
    for (size_t i = 0; i < full.size(); i++) {
      bool all_ff = true, all_zero = true;
      for (int b = 0; b < 16; b++) {
        const uint8_t x = full[i][b] ^ full1[i][b];
        all_ff &= x == 0xFF;
        all_zero &= x == 0;
      }
      CHECK(i == 1234 ? all_ff : all_zero, "two-party XOR at %zu", i);
    }
    const auto sub = scanner.eval_subtree(kp.first, 1024, 512);
    CHECK(std::memcmp(sub.data(), full.data() + 1024, 512 * 16) == 0, "subtree != slice");
    for (uint64_t i : {0ull, 1ull, 1234ull, 4095ull}) {
      uint32_t r[4];
      orc_eval(&ok, i, r);
      CHECK(std::memcmp(scanner.eval(kp.first, i).data(), r, 16) == 0, "eval(%llu)",
            (unsigned long long)i);
    }
    CHECK(throws<InvalidKeyFormat>([&] { scanner.eval_subtree(kp.first, 8, 16); }),
          "misaligned subtree");
  }

  // Errors the reference reports (or panics on).
  std::printf("case errors\n");
  std::fflush(stdout);
  CHECK(throws<RowsError>([&] { GpuPageMatrix::create(scanner, flat.data(), 100); }),
        "misaligned page data");
  CHECK(throws<RowsError>([&] {
          const auto small = make_test_pages(300);  // 300 pages > 2^8
          GpuPageMatrix big = GpuPageMatrix::create(scanner, small.data(), small.size());
          scanner.scan(big, {&queries[0][0], &queries[0][1], &queries[0][2]});
        }),
        "more pages than the domain");
  const auto k9 = generate_chacha_dpf_keys(ChaChaParams::create(9), 3, 7).first;
  CHECK(throws<InvalidDomainBits>(
            [&] { scanner.scan(db, {&queries[0][0], &k9, &queries[0][2]}); }),
        "keys of one call disagree on d");
  GpuPageMatrix empty = GpuPageMatrix::alloc_empty(scanner, 16);
  CHECK(empty.num_pages() == 16 && empty.size_bytes() == 16 * PAGE_SIZE_BYTES, "alloc_empty");
  empty.update_pages(3, flat.data(), 2 * PAGE_SIZE_BYTES);
  CHECK(throws<RowsError>([&] { empty.update_pages(15, flat.data(), 2 * PAGE_SIZE_BYTES); }),
        "update past the end");
}

void cpu_tests() {
  std::printf("case parameters and keys\n");
  CHECK(throws<InvalidDomainBits>([] { ChaChaParams::create(0); }), "d = 0");
  CHECK(throws<InvalidDomainBits>([] { ChaChaParams::create(33); }), "d = 33");
  CHECK(ChaChaParams::create(25).max_pages() == (1ull << 25), "max_pages");
  const ChaChaParams p = ChaChaParams::create(20);
  CHECK(throws<RowsError>([&] { generate_chacha_dpf_keys(p, 1u << 20, 1); }),
        "target beyond the domain (PageIndexTooLarge)");
  for (uint64_t seed : {1ull, 77ull, 123456789ull}) {
    const uint64_t target = (4242 + seed) % (1u << 20);
    const auto kp = generate_chacha_dpf_keys(p, target, seed);
    orc_key o0, o1;
    CHECK(orc_keygen(20, target, seed, &o0, &o1) == ORC_OK, "oracle keygen");
    std::vector<uint8_t> w0(orc_key_wire_len(20)), w1(orc_key_wire_len(20));
    orc_key_to_bytes(&o0, w0.data());
    orc_key_to_bytes(&o1, w1.data());
    CHECK(kp.first.to_bytes() == w0 && kp.second.to_bytes() == w1, "keygen == oracle, seed %llu",
          (unsigned long long)seed);
    CHECK(kp.first.party() == 0 && kp.second.party() == 1 && kp.first.domain_bits() == 20,
          "key header");
  }
  const auto w = generate_chacha_dpf_keys(p, 5, 9).first.to_bytes();
  CHECK(throws<InvalidKeyFormat>([&] { ChaChaKey::from_bytes(w.data(), w.size() - 1); }),
        "short key");
  std::vector<uint8_t> d33(19 + 18 * 33 + 16);
  d33[18] = 33;
  CHECK(throws<InvalidDomainBits>([&] { ChaChaKey::from_bytes(d33); }), "d = 33 key");
  CHECK(ChaChaKey::from_bytes(w).to_bytes() == w, "wire round trip");
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "--gpu";
  try {
    if (mode == "--cpu") {
      cpu_tests();
      // Without a GPU the scanner reports it instead of falling back.
      bool opened = false;
      try {
        GpuScanner s(0);
        opened = true;
      } catch (const NoDevice&) {
      } catch (const std::exception& e) {
        CHECK(false, "GpuScanner without a GPU threw %s", e.what());
      }
      if (opened) std::printf("a GPU is present: the no-device check does not apply\n");
    } else if (mode == "--gpu") {
      cpu_tests();
      gpu_tests();
    } else {
      std::fprintf(stderr, "usage: %s --cpu|--gpu\n", argv[0]);
      return 2;
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "FAIL: uncaught exception: %s\n", e.what());
    return 1;
  }
  std::printf("%s: %d failure(s)\n", mode.c_str(), g_failures);
  return g_failures ? 1 : 0;
}
