// Randomised robustness test of the wire parser/formatter and the client
// functions, built with the address and undefined-behaviour sanitizers
// (tests/cpp/Makefile target fuzz_wire_asan). Inputs are mutations of valid
// requests (bit flips, truncation, splices, duplication), random bytes,
// deeply nested and oversized documents, escapes and surrogates. Every call
// must return one of the documented statuses without touching memory outside
// its buffers; a sanitizer report aborts the run.
//
//   fuzz_wire_asan <iterations> <seed>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/morphogen_hip.h"

namespace {

std::string hex(const std::vector<uint8_t>& b) {
  static const char* d = "0123456789abcdef";
  std::string s = "0x";
  for (uint8_t c : b) {
    s += d[c >> 4];
    s += d[c & 15];
  }
  return s;
}

std::vector<uint8_t> key(uint32_t d, uint64_t seed) {
  std::vector<uint8_t> k0(19 + 18 * d + 16), k1(k0.size());
  if (mg_keygen(d, seed % (1ull << d), seed, k0.data(), k1.data()) != MG_OK) std::abort();
  return k0;
}

std::string valid_single(std::mt19937_64& rng) {
  const uint32_t d = 1 + rng() % 12;
  std::string s = "{\"keys\":[";
  for (int i = 0; i < 3; i++) s += (i ? ",\"" : "\"") + hex(key(d, rng())) + "\"";
  return s + "]}";
}

std::string valid_batch(std::mt19937_64& rng) {
  const int n = 1 + rng() % 5;
  std::string s = "{\"queries\":[";
  for (int q = 0; q < n; q++) s += (q ? "," : "") + valid_single(rng);
  return s + "]}";
}

std::string mutate(std::string s, std::mt19937_64& rng) {
  const int n = 1 + rng() % 6;
  static const char* pieces[] = {"\\u0000", "\\ud800", "\\udc00\\ud800", "\\ud83d\\ude00", "\"",
                                 "\\",      "[",      "]",             "{",           "}",
                                 ",",       ":",      "null",          "1e999",       "-0.0",
                                 "\xff",    "\xc3(",  "\xe2\x82",      "\"keys\"",    "0x"};
  for (int i = 0; i < n && !s.empty(); i++) {
    const size_t pos = rng() % (s.size() + 1);
    switch (rng() % 6) {
      case 0: s[pos % s.size()] ^= (char)(1u << (rng() % 8)); break;   // bit flip
      case 1: s.resize(pos); break;                                      // truncate
      case 2: s.insert(pos, pieces[rng() % 20]); break;                  // splice token
      case 3: s.erase(pos % s.size(), 1 + rng() % 8); break;             // delete
      case 4: s.insert(pos, s.substr(rng() % s.size(), rng() % 64)); break;  // duplicate
      case 5: s.insert(pos, 1, (char)(rng() & 0xff)); break;             // random byte
    }
  }
  return s;
}

std::string nested(int depth, bool arrays) {
  std::string s;
  for (int i = 0; i < depth; i++) s += arrays ? "[" : "{\"a\":";
  s += "1";
  for (int i = 0; i < depth; i++) s += arrays ? "]" : "}";
  return s;
}

bool allowed(int st) {
  return st == 200 || st == 400 || st == 404 || st == 405 || st == 413 || st == 415 || st == 422;
}

int g_fail = 0;

void parse_one(const std::string& body, bool batch, std::mt19937_64& rng) {
  mg_wire_opts o{};
  o.max_body_bytes = (rng() % 4) ? 0 : (1u << 20);
  o.max_batch = (rng() % 4) ? 0 : 1 + rng() % 40;
  // Exactly-sized heap buffers so the sanitizer sees any write past them.
  const size_t keys_cap = rng() % 2 ? body.size() : rng() % 4096;
  const uint32_t lens_cap = rng() % 2 ? 96 : (uint32_t)(rng() % 200);
  std::vector<uint8_t> keys(keys_cap);
  std::vector<size_t> lens(lens_cap);
  std::vector<char> buf(body.begin(), body.end());  // no terminator after the body
  uint32_t nq = 0;
  int st = 0;
  const char* method = rng() % 16 ? "POST" : "GET";
  const char* path = batch ? "/query/page/gpu/batch" : "/query/page/gpu";
  const char* ct = rng() % 16 ? "application/json" : (rng() % 2 ? nullptr : "text/plain");
  const int rc = mg_wire_parse(&o, method, path, ct, buf.data(), buf.size(),
                               keys.empty() ? nullptr : keys.data(), keys_cap,
                               lens.empty() ? nullptr : lens.data(), lens_cap, &nq, &st);
  if (!allowed(st) && rc != MG_ERR_INVALID_ARG) {
    std::fprintf(stderr, "FAIL status %d rc %d\n", st, rc);
    g_fail++;
  }
  if (rc == MG_OK) {
    if (st != 200 || nq == 0 || 3ull * nq > lens_cap) {
      std::fprintf(stderr, "FAIL ok with status %d, %u queries\n", st, nq);
      g_fail++;
    }
    size_t total = 0;
    for (uint32_t i = 0; i < 3 * nq; i++) {
      // Every returned key obeys the wire length rule 19 + 18d + 16, d <= 32.
      const uint8_t* k = keys.data() + total;
      total += lens[i];
      if (total > keys_cap || lens[i] < 35 || k[18] > 32 || lens[i] != 35 + 18 * (size_t)k[18]) {
        std::fprintf(stderr, "FAIL a returned key breaks the wire rule\n");
        g_fail++;
        break;
      }
    }
    if (total > keys_cap) g_fail++;
  }
}

void format_one(std::mt19937_64& rng) {
  const uint32_t nq = 1 + rng() % 6;
  const uint32_t w = 16 * (1 + rng() % 8);
  const int batch = nq > 1 ? 1 : (int)(rng() % 2);
  std::vector<uint8_t> pages(3ull * nq * w);
  for (auto& b : pages) b = (uint8_t)rng();
  const size_t need = mg_wire_response_capacity(nq, w);
  const size_t cap = rng() % 3 ? need : rng() % need;
  std::vector<char> out(cap + 1);
  size_t len = 0;
  const int rc = mg_wire_format(rng(), pages.data(), nq, w, batch, out.data(), cap, &len);
  if (rc == MG_OK && (cap < need || len >= cap)) {
    std::fprintf(stderr, "FAIL format wrote %zu into %zu\n", len, cap);
    g_fail++;
  }
  // The body is {"epoch_id":E,"pages":[...]} or {"epoch_id":E,"results":[...]}
  // with 3 x nq pages of "0x" + 2W hex characters each.
  if (rc == MG_OK) {
    const std::string body(out.data(), len);
    const size_t hex_chars = 3ull * nq * (2 + 2 * (size_t)w);
    if (body.rfind("{\"epoch_id\":", 0) != 0 || body.back() != '}' || len < hex_chars ||
        body.find(batch ? "\"results\":[" : "\"pages\":[") == std::string::npos) {
      std::fprintf(stderr, "FAIL format shape\n");
      g_fail++;
    }
  }
}

void client_one(std::mt19937_64& rng) {
  std::vector<uint8_t> acct(rng() % 80);
  for (auto& b : acct) b = (uint8_t)rng();
  const uint64_t seeds[3] = {rng(), rng(), rng()};
  uint64_t out[3];
  const uint64_t rows = rng() % 3 ? 1 + rng() % (1ull << 40) : 0;
  const int rc = mg_cuckoo_hash_indices(acct.empty() ? nullptr : acct.data(), acct.size(), rows,
                                        seeds, out);
  if (rc == MG_OK)
    for (int i = 0; i < 3; i++)
      if (out[i] >= rows) g_fail++;
  const uint32_t d = rng() % 34;
  const uint64_t pages = rng() % 5 ? 1 + rng() % (1ull << (d < 20 ? d : 20)) : rng();
  const size_t klen = 19 + 18 * (size_t)(d <= 32 ? d : 0) + 16;
  std::vector<uint8_t> ka(3 * klen), kb(3 * klen);
  uint64_t pidx[3];
  uint32_t offs[3];
  if (mg_page_query(acct.empty() ? nullptr : acct.data(), acct.size(), pages, seeds, d, rng(),
                    ka.data(), kb.data(), pidx, offs) == MG_OK)
    for (int i = 0; i < 3; i++)
      if (pidx[i] >= pages || offs[i] >= MG_ROWS_PER_PAGE) g_fail++;
  std::vector<uint8_t> page(rng() % 5000), row(MG_PAGE_ROW_BYTES);
  mg_page_extract_row(page.empty() ? nullptr : page.data(), page.size(), (uint32_t)(rng() % 20),
                      row.data());
}

}  // namespace

int main(int argc, char** argv) {
  const long iters = argc > 1 ? std::atol(argv[1]) : 20000;
  std::mt19937_64 rng(argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1);
  for (long it = 0; it < iters; it++) {
    std::string body;
    bool batch = rng() % 2;
    switch (rng() % 8) {
      case 0: body = batch ? valid_batch(rng) : valid_single(rng); break;
      case 1:
      case 2:
      case 3: body = mutate(batch ? valid_batch(rng) : valid_single(rng), rng); break;
      case 4: {
        body.resize(rng() % 300);
        for (auto& c : body) c = (char)rng();
        break;
      }
      case 5: body = nested(1 + rng() % 300, rng() % 2); break;
      case 6: body = "{\"keys\":[\"" + std::string(rng() % 70000, 'a') + "\"]}"; break;
      case 7: body = mutate(nested(1 + rng() % 140, rng() % 2), rng); break;
    }
    parse_one(body, batch, rng);
    if (it % 8 == 0) format_one(rng);
    if (it % 4 == 0) client_one(rng);
  }
  std::printf("fuzz: %ld iterations, %d failure(s)\n", iters, g_fail);
  return g_fail ? 1 : 0;
}
