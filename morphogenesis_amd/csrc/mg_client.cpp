// Client side (host code): DPF key generation, cuckoo addressing, page
// queries and the reconstruction of answers, as the reference client does
// (crates/morphogen-client/src/gpu.rs:17-49 and page.rs:67-163 over
// crates/morphogen-gpu-dpf/src/dpf.rs:322-432 and
// crates/morphogen-core/src/cuckoo.rs:30-40).
// Root seeds come from a caller-given 64-bit seed instead of the OS RNG so
// load generators and tests are reproducible; keys are emitted in the wire
// format of dpf.rs:259-274.
#include <stdint.h>
#include <string.h>

#include "../../include/morphogen_hip.h"

namespace {

inline uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

// ChaCha8 block over key = seed||seed, nonce 0 (chacha_prg.rs:112-178).
void chacha8_seed_block(const uint32_t s[4], uint32_t counter, uint32_t out[16]) {
  const uint32_t init[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                             s[0], s[1], s[2], s[3], s[0], s[1], s[2], s[3],
                             counter, 0, 0, 0};
  uint32_t x[16];
  memcpy(x, init, sizeof x);
  auto qr = [&](int a, int b, int c, int d) {
    x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 16);
    x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 12);
    x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl(x[d], 8);
    x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl(x[b], 7);
  };
  for (int r = 0; r < 4; r++) {
    qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
    qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + init[i];
}

struct Children {
  uint32_t l[4], r[4];
  uint8_t lt, rt;
};

Children expand(const uint32_t s[4]) {
  uint32_t b0[16], b1[16];
  chacha8_seed_block(s, 0, b0);
  chacha8_seed_block(s, 1, b1);
  Children c;
  memcpy(c.l, b0, 16);
  memcpy(c.r, b0 + 4, 16);
  c.lt = (uint8_t)(b1[0] & 1u);
  c.rt = (uint8_t)(b1[1] & 1u);
  return c;
}

uint64_t splitmix64(uint64_t seed, uint64_t j) {
  uint64_t z = seed + (j + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void put_seed(uint8_t* p, const uint32_t s[4]) {
  for (int i = 0; i < 4; i++)
    for (int b = 0; b < 4; b++) p[4 * i + b] = (uint8_t)(s[i] >> (8 * b));
}

}  // namespace

extern "C" int mg_keygen(uint32_t domain_bits, uint64_t target, uint64_t seed, uint8_t* key0,
                         uint8_t* key1) {
  if (!key0 || !key1) return MG_ERR_INVALID_ARG;
  if (domain_bits == 0 || domain_bits > 32) return MG_ERR_DOMAIN;
  if (target >> domain_bits) return MG_ERR_ROWS;
  const uint32_t d = domain_bits;
  uint32_t root0[4], root1[4];
  for (int i = 0; i < 2; i++) {
    const uint64_t a = splitmix64(seed, (uint64_t)i), b = splitmix64(seed, (uint64_t)(2 + i));
    root0[2 * i] = (uint32_t)a;
    root0[2 * i + 1] = (uint32_t)(a >> 32);
    root1[2 * i] = (uint32_t)b;
    root1[2 * i + 1] = (uint32_t)(b >> 32);
  }
  uint32_t s0[4], s1[4];
  memcpy(s0, root0, 16);
  memcpy(s1, root1, 16);
  uint8_t t0 = 0, t1 = 1;
  // Header and correction words are shared; only root seed / t / party differ.
  size_t off = 19;
  for (uint32_t level = 0; level < d; level++) {
    const Children c0 = expand(s0), c1 = expand(s1);
    const int bit = (int)((target >> (d - 1 - level)) & 1u);
    const uint32_t* keep0 = bit ? c0.r : c0.l;
    const uint32_t* lose0 = bit ? c0.l : c0.r;
    const uint32_t* keep1 = bit ? c1.r : c1.l;
    const uint32_t* lose1 = bit ? c1.l : c1.r;
    const uint8_t keep_t0 = bit ? c0.rt : c0.lt, lose_t0 = bit ? c0.lt : c0.rt;
    const uint8_t keep_t1 = bit ? c1.rt : c1.lt, lose_t1 = bit ? c1.lt : c1.rt;
    uint32_t cw[4];
    for (int i = 0; i < 4; i++) cw[i] = lose0[i] ^ lose1[i];
    const uint8_t t_cw_lose = (uint8_t)(lose_t0 ^ lose_t1 ^ t0 ^ t1 ^ 1);
    const uint8_t t_cw_keep = (uint8_t)(keep_t0 ^ keep_t1 ^ t0 ^ t1);
    put_seed(key0 + off, cw);
    key0[off + 16] = bit ? t_cw_lose : t_cw_keep;
    key0[off + 17] = bit ? t_cw_keep : t_cw_lose;
    off += 18;
    uint32_t n0[4], n1[4];
    for (int i = 0; i < 4; i++) {
      n0[i] = keep0[i] ^ (t0 == 1 ? cw[i] : 0u);
      n1[i] = keep1[i] ^ (t1 == 1 ? cw[i] : 0u);
    }
    const uint8_t nt0 = (uint8_t)(keep_t0 ^ (t0 == 1 ? t_cw_keep : 0));
    const uint8_t nt1 = (uint8_t)(keep_t1 ^ (t1 == 1 ? t_cw_keep : 0));
    memcpy(s0, n0, 16);
    memcpy(s1, n1, 16);
    t0 = nt0;
    t1 = nt1;
  }
  uint32_t fcw[4];
  for (int i = 0; i < 4; i++) fcw[i] = s0[i] ^ s1[i] ^ 0xFFFFFFFFu;
  put_seed(key0 + off, fcw);
  memcpy(key1 + 19, key0 + 19, off + 16 - 19);
  put_seed(key0, root0);
  put_seed(key1, root1);
  key0[16] = 0;  // root_t
  key0[17] = 0;  // party
  key0[18] = (uint8_t)d;
  key1[16] = 1;
  key1[17] = 1;
  key1[18] = (uint8_t)d;
  return MG_OK;
}

// ------------------------------------------------------ cuckoo addressing --
namespace {

// SipHash-1-3 with the all-zero key: what Rust's std DefaultHasher::new()
// computes, over the bytes written to it. Written from the published
// SipHash construction (Aumasson & Bernstein): four 64-bit lanes, one
// SipRound per 8-byte little-endian word, the length mod 256 in the top byte
// of the last word, three finalisation rounds after v2 ^= 0xff.
struct Sip13 {
  uint64_t v0 = 0x736f6d6570736575ull, v1 = 0x646f72616e646f6dull;
  uint64_t v2 = 0x6c7967656e657261ull, v3 = 0x7465646279746573ull;
  static uint64_t rotl(uint64_t x, int b) { return (x << b) | (x >> (64 - b)); }
  void round() {
    v0 += v1; v1 = rotl(v1, 13); v1 ^= v0; v0 = rotl(v0, 32);
    v2 += v3; v3 = rotl(v3, 16); v3 ^= v2;
    v0 += v3; v3 = rotl(v3, 21); v3 ^= v0;
    v2 += v1; v1 = rotl(v1, 17); v1 ^= v2; v2 = rotl(v2, 32);
  }
  void absorb(uint64_t m) {
    v3 ^= m;
    round();
    v0 ^= m;
  }
  uint64_t finish(uint64_t last) {
    absorb(last);
    v2 ^= 0xff;
    round();
    round();
    round();
    return v0 ^ v1 ^ v2 ^ v3;
  }
};

// DefaultHasher: write_u64(seed) then write(key) (cuckoo.rs:33-37). The seed
// is exactly one 8-byte word, so the key's words start word-aligned.
uint64_t cuckoo_hash(uint64_t seed, const uint8_t* key, size_t len) {
  Sip13 h;
  h.absorb(seed);
  size_t i = 0;
  for (; i + 8 <= len; i += 8) {
    uint64_t m = 0;
    for (int b = 0; b < 8; b++) m |= (uint64_t)key[i + b] << (8 * b);
    h.absorb(m);
  }
  uint64_t last = (uint64_t)((len + 8) & 0xff) << 56;
  for (size_t b = 0; i + b < len; b++) last |= (uint64_t)key[i + b] << (8 * b);
  return h.finish(last);
}

}  // namespace

extern "C" int mg_cuckoo_hash_indices(const uint8_t* key, size_t key_len, uint64_t num_rows,
                                      const uint64_t seeds[3], uint64_t out[3]) {
  if ((!key && key_len) || !seeds || !out) return MG_ERR_INVALID_ARG;
  if (num_rows == 0) return MG_ERR_ROWS;
  for (int i = 0; i < 3; i++) out[i] = cuckoo_hash(seeds[i], key, key_len) % num_rows;
  return MG_OK;
}

extern "C" int mg_page_query(const uint8_t* account_key, size_t key_len, uint64_t num_pages,
                             const uint64_t seeds[3], uint32_t domain_bits, uint64_t rng_seed,
                             uint8_t* keys_a, uint8_t* keys_b, uint64_t page_index[3],
                             uint32_t row_offset[3]) {
  if (!keys_a || !keys_b || !page_index || !row_offset) return MG_ERR_INVALID_ARG;
  if (domain_bits == 0 || domain_bits > 32) return MG_ERR_DOMAIN;
  // num_pages * ROWS_PER_PAGE must not overflow (gpu.rs:23-29).
  if (num_pages > UINT64_MAX / MG_ROWS_PER_PAGE) return MG_ERR_DOMAIN;
  uint64_t rows[3];
  int rc = mg_cuckoo_hash_indices(account_key, key_len, num_pages * MG_ROWS_PER_PAGE, seeds, rows);
  if (rc) return rc;
  const size_t klen = 19 + 18 * (size_t)domain_bits + 16;
  for (int i = 0; i < 3; i++) {
    page_index[i] = rows[i] / MG_ROWS_PER_PAGE;  // PageAddress::from_row_index
    row_offset[i] = (uint32_t)(rows[i] % MG_ROWS_PER_PAGE);
    // Independent root seeds per key pair.
    rc = mg_keygen(domain_bits, page_index[i], splitmix64(rng_seed, 16 + (uint64_t)i),
                   keys_a + i * klen, keys_b + i * klen);
    if (rc) return rc;  // MG_ERR_ROWS: page index beyond 2^d (PageIndexTooLarge)
  }
  return MG_OK;
}

extern "C" int mg_page_aggregate(uint64_t epoch_a, const uint8_t* pages_a, uint64_t epoch_b,
                                 const uint8_t* pages_b, size_t page_bytes, uint32_t n_pages,
                                 uint8_t* out) {
  if ((!pages_a || !pages_b || !out) && page_bytes && n_pages) return MG_ERR_INVALID_ARG;
  if (epoch_a != epoch_b) return MG_ERR_STATE;  // PageAggregationError::EpochMismatch
  const size_t n = page_bytes * n_pages;
  for (size_t i = 0; i < n; i++) out[i] = (uint8_t)(pages_a[i] ^ pages_b[i]);
  return MG_OK;
}

extern "C" int mg_page_extract_row(const uint8_t* page, size_t page_bytes, uint32_t row_offset,
                                   uint8_t* out) {
  if (!page || !out) return MG_ERR_INVALID_ARG;
  if (row_offset >= MG_ROWS_PER_PAGE) return MG_ERR_ROWS;  // ExtractError::InvalidRowOffset
  const size_t start = (size_t)row_offset * MG_PAGE_ROW_BYTES;
  if (start + MG_PAGE_ROW_BYTES > page_bytes) return MG_ERR_ROWS;
  memcpy(out, page + start, MG_PAGE_ROW_BYTES);
  return MG_OK;
}

