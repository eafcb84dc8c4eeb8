/*
 * morphogen_hip.hpp - C++17 host interface over the C ABI (morphogen_hip.h),
 * shaped like the reference's Rust API for this path so a C++ host reads like
 * the reference's own call sites and tests:
 *   ChaChaParams, ChaChaKey::{from_bytes,to_bytes}, generate_chacha_dpf_keys
 *     (crates/morphogen-gpu-dpf/src/dpf.rs:60-99, 259-432)
 *   GpuScanner::{new,scan,scan_batch,scan_batch_optimized(_tiled)(_with_graph),
 *     scan_batch_single_query_multistream_optimized} (kernel.rs:158-171, 258,
 *     685-1162), and ChaChaKey::{eval,eval_subtree,full_eval} (dpf.rs:103-246)
 *     as GpuScanner methods, since they run on the scanner's device
 *   GpuPageMatrix::{new,alloc_empty,update_pages,num_pages,size_bytes}
 *     (storage.rs:18-145)
 *   PirResult, KernelTiming (kernel.rs:1213-1240)
 * Header only; every call goes to the GPU through libmorphogen_hip.so. Errors
 * are thrown as MgError (or a subclass named after GpuDpfError's variants,
 * dpf.rs:9-40) where the reference returns Err; the matrix constructors throw
 * RowsError where GpuPageMatrix panics (storage.rs:31-33, 109-116).
 *
 * Differences, stated: keys are seeded (generate_chacha_dpf_keys takes the
 * seed its root seeds are derived from; the reference draws them from the OS
 * RNG), a page is one row of the matrix's row width (PAGE_SIZE_BYTES by
 * default), and domains go up to 32 bits.
 */
#ifndef MORPHOGEN_HIP_HPP
#define MORPHOGEN_HIP_HPP

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "morphogen_hip.h"

namespace morphogen {

constexpr size_t PAGE_SIZE_BYTES = MG_PAGE_BYTES;  // storage.rs PAGE_SIZE_BYTES
constexpr size_t KEYS_PER_QUERY = 3;               // kernel.rs KEYS_PER_QUERY
constexpr uint32_t MAX_DOMAIN_BITS = 32;           // the kernel's key record

class MgError : public std::runtime_error {
 public:
  MgError(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};
struct InvalidKeyFormat : MgError { using MgError::MgError; };
struct InvalidDomainBits : MgError { using MgError::MgError; };
struct RowsError : MgError { using MgError::MgError; };
struct NoDevice : MgError { using MgError::MgError; };

inline void check(int rc, const char* what) {
  if (rc == MG_OK) return;
  std::string msg = std::string(what) + ": " + mg_last_error();
  switch (rc) {
    case MG_ERR_KEY_FORMAT: throw InvalidKeyFormat(rc, msg);
    case MG_ERR_DOMAIN: throw InvalidDomainBits(rc, msg);
    case MG_ERR_ROWS: throw RowsError(rc, msg);
    case MG_ERR_NO_DEVICE: throw NoDevice(rc, msg);
    default: throw MgError(rc, msg);
  }
}

// ChaChaParams (dpf.rs:60-90).
struct ChaChaParams {
  uint32_t domain_bits;
  static ChaChaParams create(uint32_t domain_bits) {
    if (domain_bits == 0 || domain_bits > MAX_DOMAIN_BITS)
      throw InvalidDomainBits(MG_ERR_DOMAIN, "domain_bits must be in 1..32");
    return ChaChaParams{domain_bits};
  }
  uint64_t max_pages() const { return 1ull << domain_bits; }
};

// ChaChaKey kept in its wire form (dpf.rs:259-316): what crosses the ABI.
class ChaChaKey {
 public:
  static ChaChaKey from_bytes(const uint8_t* bytes, size_t len) {
    mg_key_info info{};
    check(mg_key_parse(bytes, len, &info), "ChaChaKey::from_bytes");
    ChaChaKey k;
    k.wire_.assign(bytes, bytes + len);
    k.info_ = info;
    return k;
  }
  static ChaChaKey from_bytes(const std::vector<uint8_t>& b) { return from_bytes(b.data(), b.size()); }
  const std::vector<uint8_t>& to_bytes() const { return wire_; }
  uint32_t domain_bits() const { return info_.domain_bits; }
  uint32_t party() const { return info_.party; }

 private:
  std::vector<uint8_t> wire_;
  mg_key_info info_{};
};

// generate_chacha_dpf_keys (dpf.rs:322-432) with seeded root seeds.
inline std::pair<ChaChaKey, ChaChaKey> generate_chacha_dpf_keys(const ChaChaParams& p,
                                                                uint64_t target, uint64_t seed) {
  const size_t n = 19 + 18 * (size_t)p.domain_bits + 16;
  std::vector<uint8_t> k0(n), k1(n);
  check(mg_keygen(p.domain_bits, target, seed, k0.data(), k1.data()), "generate_chacha_dpf_keys");
  return {ChaChaKey::from_bytes(k0), ChaChaKey::from_bytes(k1)};
}

// KernelTiming (kernel.rs:1213-1228), nanoseconds, per call.
struct KernelTiming {
  uint64_t h2d_ns = 0, kernel_ns = 0, d2h_ns = 0, total_ns = 0;
  static KernelTiming from(const mg_timing& t) {
    auto ns = [](double ms) { return (uint64_t)(ms * 1e6); };
    return {ns(t.h2d_ms), ns(t.kernel_ms), ns(t.d2h_ms), ns(t.total_ms)};
  }
};

// PirResult (kernel.rs:1230-1240): three answers; the verifier words are
// zero, as in the reference build with verifiable PIR compiled out.
struct PirResult {
  std::vector<uint8_t> page0, page1, page2;
  std::vector<uint8_t> verif0 = std::vector<uint8_t>(16), verif1 = std::vector<uint8_t>(16),
                       verif2 = std::vector<uint8_t>(16);
  KernelTiming timing;
};

class GpuScanner;

// GpuPageMatrix (storage.rs:18-145): device rows of row_bytes each.
class GpuPageMatrix {
 public:
  GpuPageMatrix(const GpuPageMatrix&) = delete;
  GpuPageMatrix& operator=(const GpuPageMatrix&) = delete;
  GpuPageMatrix(GpuPageMatrix&& o) noexcept : m_(o.m_), rows_(o.rows_), w_(o.w_) { o.m_ = nullptr; }
  ~GpuPageMatrix() {
    if (m_) mg_matrix_destroy(m_);
  }

  // GpuPageMatrix::new (storage.rs:27-43): data must be whole pages.
  static GpuPageMatrix create(const GpuScanner& s, const uint8_t* data, size_t bytes,
                              uint32_t row_bytes = PAGE_SIZE_BYTES);
  // GpuPageMatrix::alloc_empty (storage.rs:89-101).
  static GpuPageMatrix alloc_empty(const GpuScanner& s, uint64_t num_pages,
                                   uint32_t row_bytes = PAGE_SIZE_BYTES);
  // GpuPageMatrix::update_pages (storage.rs:104-124).
  void update_pages(uint64_t start_page, const uint8_t* data, size_t bytes) {
    if (bytes % w_) throw RowsError(MG_ERR_ROWS, "update is not a whole number of pages");
    check(mg_matrix_upload(m_, start_page, data, bytes / w_), "update_pages");
  }
  uint64_t num_pages() const { return rows_; }
  uint64_t size_bytes() const { return rows_ * w_; }
  uint32_t row_bytes() const { return w_; }
  mg_matrix* raw() const { return m_; }

 private:
  GpuPageMatrix(mg_matrix* m, uint64_t rows, uint32_t w) : m_(m), rows_(rows), w_(w) {}
  mg_matrix* m_ = nullptr;
  uint64_t rows_ = 0;
  uint32_t w_ = 0;
};

// GpuScanner (kernel.rs:158-171): one device. Safe to share across threads;
// scans from several threads run concurrently on the device's workspaces.
class GpuScanner {
 public:
  explicit GpuScanner(int device_ord = 0) { check(mg_device_open(device_ord, &dev_), "GpuScanner::new"); }
  GpuScanner(const GpuScanner&) = delete;
  GpuScanner& operator=(const GpuScanner&) = delete;
  ~GpuScanner() {
    if (dev_) mg_device_close(dev_);
  }
  mg_device* raw() const { return dev_; }
  // Workspace slots (streams) the device's calls spread over, 1-8 (kernel.rs:50-65).
  void set_workspaces(uint32_t n) { check(mg_device_set_workspaces(dev_, n), "set_workspaces"); }

  using Seed128 = std::array<uint8_t, 16>;  // a leaf: four little-endian words
  // ChaChaKey::eval (dpf.rs:103-142), on this device.
  Seed128 eval(const ChaChaKey& k, uint64_t index) const {
    Seed128 out{};
    check(mg_key_eval(dev_, k.to_bytes().data(), k.to_bytes().size(), &index, 1, out.data()),
          "ChaChaKey::eval");
    return out;
  }
  // ChaChaKey::eval_subtree (dpf.rs:160-196): leaves [start, start + size).
  std::vector<Seed128> eval_subtree(const ChaChaKey& k, uint64_t start, uint64_t size) const {
    std::vector<Seed128> out(size);
    check(mg_key_eval_subtree(dev_, k.to_bytes().data(), k.to_bytes().size(), start, size,
                              out.empty() ? nullptr : out.front().data()),
          "ChaChaKey::eval_subtree");
    return out;
  }
  // ChaChaKey::full_eval (dpf.rs:147-158): all 2^d leaves.
  std::vector<Seed128> full_eval(const ChaChaKey& k) const {
    std::vector<Seed128> out(1ull << k.domain_bits());
    check(mg_key_full_eval(dev_, k.to_bytes().data(), k.to_bytes().size(), out.front().data(),
                           out.size() * 16),
          "ChaChaKey::full_eval");
    return out;
  }

  // scan (kernel.rs:685-709): one query of three keys in one pass.
  PirResult scan(const GpuPageMatrix& db, const std::array<const ChaChaKey*, 3>& keys) const {
    std::vector<std::array<const ChaChaKey*, 3>> q{keys};
    return std::move(run(db, q, mg_scan_opts{}, false).front());
  }
  // scan_batch (kernel.rs:711-924): every query's keys in one pass.
  std::vector<PirResult> scan_batch(const GpuPageMatrix& db,
                                    const std::vector<std::array<ChaChaKey, 3>>& queries) const {
    return run(db, refs(queries), mg_scan_opts{}, false);
  }
  std::vector<PirResult> scan_batch_optimized(
      const GpuPageMatrix& db, const std::vector<std::array<ChaChaKey, 3>>& queries) const {
    return scan_batch(db, queries);
  }
  std::vector<PirResult> scan_batch_optimized_with_graph(
      const GpuPageMatrix& db, const std::vector<std::array<ChaChaKey, 3>>& queries,
      bool use_graph) const {
    mg_scan_opts o{};
    o.use_graph = use_graph ? 1 : 0;
    return run(db, refs(queries), o, false);
  }
  // scan_batch_optimized_tiled(_with_graph) (kernel.rs:1099-1162): the
  // reference launches queries in tiles of tile_size; here every tile is part
  // of one pass, so the tile size does not change the answers.
  std::vector<PirResult> scan_batch_optimized_tiled(
      const GpuPageMatrix& db, const std::vector<std::array<ChaChaKey, 3>>& queries,
      size_t tile_size, bool use_graph = false) const {
    if (tile_size == 0) throw MgError(MG_ERR_INVALID_ARG, "tile_size must be positive");
    return scan_batch_optimized_with_graph(db, queries, use_graph);
  }
  // scan_batch_single_query_multistream_optimized(_with_graph)
  // (kernel.rs:955-1097): one query per launch.
  std::vector<PirResult> scan_batch_single_query_multistream_optimized(
      const GpuPageMatrix& db, const std::vector<std::array<ChaChaKey, 3>>& queries,
      size_t stream_count, bool use_graph = false) const {
    if (stream_count == 0) throw MgError(MG_ERR_INVALID_ARG, "stream_count must be positive");
    mg_scan_opts o{};
    o.use_graph = use_graph ? 1 : 0;
    return run(db, refs(queries), o, true);
  }

 private:
  using QueryRefs = std::vector<std::array<const ChaChaKey*, 3>>;
  static QueryRefs refs(const std::vector<std::array<ChaChaKey, 3>>& queries) {
    QueryRefs r;
    for (const auto& q : queries) r.push_back({&q[0], &q[1], &q[2]});
    return r;
  }
  std::vector<PirResult> run(const GpuPageMatrix& db, const QueryRefs& queries,
                             const mg_scan_opts& o, bool per_query) const {
    std::vector<uint8_t> blob;
    std::vector<size_t> lens;
    for (const auto& q : queries)
      for (const ChaChaKey* k : q) {
        blob.insert(blob.end(), k->to_bytes().begin(), k->to_bytes().end());
        lens.push_back(k->to_bytes().size());
      }
    const uint32_t n = (uint32_t)lens.size();
    const size_t w = db.row_bytes();
    std::vector<uint8_t> out((size_t)n * w);
    mg_timing t{};
    if (n) {
      if (per_query)
        check(mg_scan_per_query(dev_, db.raw(), blob.data(), lens.data(), n, 3, out.data(), &o, &t),
              "scan_batch_single_query_multistream_optimized");
      else
        check(mg_scan(dev_, db.raw(), blob.data(), lens.data(), n, out.data(), &o, &t), "scan");
    }
    std::vector<PirResult> res(queries.size());
    for (size_t i = 0; i < queries.size(); i++) {
      const uint8_t* a = out.data() + 3 * i * w;
      res[i].page0.assign(a, a + w);
      res[i].page1.assign(a + w, a + 2 * w);
      res[i].page2.assign(a + 2 * w, a + 3 * w);
      res[i].timing = KernelTiming::from(t);
    }
    return res;
  }
  mg_device* dev_ = nullptr;
};

inline GpuPageMatrix GpuPageMatrix::create(const GpuScanner& s, const uint8_t* data, size_t bytes,
                                           uint32_t row_bytes) {
  if (row_bytes == 0 || bytes % row_bytes)
    throw RowsError(MG_ERR_ROWS, "data is not a whole number of pages");
  GpuPageMatrix m = alloc_empty(s, bytes / row_bytes, row_bytes);
  if (bytes) check(mg_matrix_upload(m.m_, 0, data, bytes / row_bytes), "GpuPageMatrix::new");
  return m;
}

inline GpuPageMatrix GpuPageMatrix::alloc_empty(const GpuScanner& s, uint64_t num_pages,
                                                uint32_t row_bytes) {
  mg_matrix* m = nullptr;
  check(mg_matrix_create(s.raw(), num_pages, row_bytes, &m), "GpuPageMatrix::alloc_empty");
  return GpuPageMatrix(m, num_pages, row_bytes);
}

}  // namespace morphogen

#endif
