/*
 * CPU oracle for the ChaCha8 DPF + masked-XOR PIR scan.
 *
 * TEST INFRASTRUCTURE ONLY. This library computes the expected answers the HIP path is
 * compared against, and the CPU baseline that is timed beside it. Nothing in
 * the product library links, loads or calls it.
 *
 * It restates, in plain C, the reference's CPU algorithm for this path
 * (igor53627/morphogenesis, Rust):
 *   - ChaCha8 block + PRG expand         crates/morphogen-gpu-dpf/src/chacha_prg.rs:92-178
 *   - ChaChaKey eval / subtree / full    crates/morphogen-gpu-dpf/src/dpf.rs:103-246
 *   - wire format to_bytes / from_bytes  crates/morphogen-gpu-dpf/src/dpf.rs:259-316
 *   - keygen                             crates/morphogen-gpu-dpf/src/dpf.rs:322-432
 *   - masked XOR + fused tiled scan      crates/morphogen-gpu-dpf/src/kernel.rs:1243-1383
 *   - delta scan                         crates/morphogen-server/src/scan.rs:434-485
 *   - byte-LCG fill                      crates/morphogen-storage/src/chunked_matrix.rs:83-93
 *   - cuckoo addressing                  crates/morphogen-core/src/cuckoo.rs:30-40 (Rust std
 *                                        DefaultHasher = SipHash-1-3, all-zero key)
 * generalised from 4096-byte pages to W-byte rows (W a multiple of 16) with one
 * DPF leaf per row.
 *
 * How each function is pinned (checks live in tests/test_oracle.py):
 *   orc_chacha_block   rounds=20 equals OpenSSL 3 chacha20 keystream (committed
 *                      fixture tests/golden/chacha20_openssl.json); the round
 *                      count is a loop bound, rounds=8 runs the same code.
 *   orc_prg_expand     rounds=20 wiring (seed||seed key, counters 0/1, word
 *                      selection, low-bit control bits) equals the OpenSSL
 *                      keystream for key = seed||seed.
 *   orc_eval etc.      two-party XOR is the point function on full domains
 *                      d<=16; full_eval == point eval (pins MSB-first order);
 *                      eval_subtree == slice of full_eval; boundary targets.
 *   orc_key_*_bytes    exact length rule 19+18d+16, round trip, field offsets
 *                      checked on hand-built byte strings.
 *   orc_scan_*         brute force (per-row eval) == tiled form; two-party
 *                      XOR == target row; delta linearity; closed forms for
 *                      all-zero / all-ones rows.
 *   oracle_fast.c      orc_eval_subtree_fast equals orc_eval_subtree on
 *                      random and fuzzed keys; orc_scan_fast, orc_scan_rows8
 *                      and orc_scan_seeded equal orc_scan_tiled (and brute
 *                      force) on the same keys and rows, whole and in ranges.
 *   orc_lcg_fill       jump-ahead equals the sequential byte loop.
 *   orc_siphash        (c, d) = (1, 3) and (2, 4) equal OpenSSL 3's SIPHASH MAC
 *                      with c-rounds/d-rounds set (committed fixture
 *                      tests/golden/siphash_openssl.json); (2, 4) with the
 *                      zero key also equals CPython 3.10's bytes hash under
 *                      PYTHONHASHSEED=0 (an independent implementation).
 *   orc_cuckoo_indices the message layout (LE64(seed) || key, then % rows) is
 *                      pinned by reading cuckoo.rs:30-40 and Rust's Hasher
 *                      (write_u64 writes the native little-endian bytes).
 * Wire-format field semantics (which byte is t_cw_left) are pinned by reading
 * dpf.rs:259-316 only: no reference-generated key exists to compare against.
 */
#ifndef MORPHOGEN_ORACLE_H
#define MORPHOGEN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_DOMAIN_BITS 32

enum {
  ORC_OK = 0,
  ORC_ERR_KEY_FORMAT = 1,   /* GpuDpfError::InvalidKeyFormat        */
  ORC_ERR_DOMAIN = 2,       /* GpuDpfError::InvalidDomainBits       */
  ORC_ERR_INDEX = 3,        /* GpuDpfError::PageIndexTooLarge       */
  ORC_ERR_SUBTREE = 4,      /* eval_subtree alignment / size errors */
  ORC_ERR_ROWS = 5,         /* more rows than 2^d, or bad row width */
};

/* In-memory ChaChaKey (dpf.rs:91-99). Seeds are 4 little-endian u32 words. */
typedef struct {
  uint32_t root_seed[4];
  uint8_t root_t;
  uint8_t party;
  uint8_t domain_bits;
  uint32_t cw_seed[ORC_MAX_DOMAIN_BITS][4];
  uint8_t cw_t_left[ORC_MAX_DOMAIN_BITS];
  uint8_t cw_t_right[ORC_MAX_DOMAIN_BITS];
  uint32_t final_cw[4];
} orc_key;

void orc_chacha_block(const uint32_t key[8], uint32_t counter,
                      const uint32_t nonce[3], int rounds, uint32_t out[16]);
void orc_prg_expand(const uint32_t seed[4], int rounds, uint32_t left[4],
                    uint32_t right[4], uint8_t *left_t, uint8_t *right_t);

size_t orc_key_wire_len(int domain_bits);
int orc_key_from_bytes(const uint8_t *bytes, size_t len, orc_key *out);
size_t orc_key_to_bytes(const orc_key *key, uint8_t *out);

void orc_eval(const orc_key *key, uint64_t index, uint32_t out[4]);
int orc_eval_subtree(const orc_key *key, uint64_t start, uint64_t size,
                     uint32_t *out);
int orc_full_eval(const orc_key *key, uint32_t *out);

int orc_keygen(int domain_bits, uint64_t target, uint64_t seed, orc_key *k0,
               orc_key *k1);

int orc_scan_brute(const orc_key *key, const uint8_t *rows, uint64_t n_rows,
                   uint32_t row_bytes, uint8_t *out);
int orc_scan_tiled(const orc_key *keys, int n_keys, const uint8_t *rows,
                   uint64_t n_rows, uint32_t row_bytes, int n_threads,
                   uint8_t *out);
int orc_scan_delta(const orc_key *keys, int n_keys, const uint64_t *row_idx,
                   const uint8_t *diffs, uint64_t n_entries,
                   uint32_t row_bytes, uint8_t *out);

/* The node on index's path at tree level `levels` (first steps of orc_eval). */
void orc_path_node(const orc_key *key, uint64_t index, int levels,
                   uint32_t seed[4], uint8_t *t);

/* oracle_fast.c: the same results computed with 16-lane vector ChaCha8 and
 * wide masked-XOR folds (pinned against the scalar functions above).
 * orc_eval_subtree_fast: size <= 2^d and <= 2^20, else ORC_ERR_SUBTREE.
 * orc_scan_fast / orc_scan_seeded: all keys share one domain (ORC_ERR_DOMAIN).
 * orc_scan_seeded: answers over rows [row_begin, row_end) of the n_rows-row
 * matrix that orc_fill_rows(seed) generates, without holding it.
 * orc_scan_rows8: the reference server scan's eight-rows-per-step shape. */
int orc_eval_subtree_fast(const orc_key *key, uint64_t start, uint64_t size,
                          uint32_t *out);
int orc_scan_fast(const orc_key *keys, int n_keys, const uint8_t *rows,
                  uint64_t n_rows, uint32_t row_bytes, int n_threads,
                  uint8_t *out);
int orc_scan_seeded(const orc_key *keys, int n_keys, uint64_t seed,
                    uint64_t n_rows, uint64_t row_begin, uint64_t row_end,
                    uint32_t row_bytes, int n_threads, uint8_t *out);
int orc_scan_rows8(const orc_key *keys, int n_keys, const uint8_t *rows,
                   uint64_t n_rows, uint32_t row_bytes, int n_threads,
                   uint8_t *out);
const char *orc_simd_isa(void);

uint64_t orc_splitmix64_word(uint64_t seed, uint64_t j);
void orc_fill_rows(uint64_t seed, uint64_t first_row, uint64_t n_rows,
                   uint32_t row_bytes, uint8_t *out);
void orc_lcg_fill(uint64_t seed, uint64_t first_byte, uint64_t n_bytes,
                  uint8_t *out);
uint64_t orc_lcg_state_at(uint64_t seed, uint64_t steps);

/* SipHash-c-d (Aumasson & Bernstein 2012) of msg under key (k0, k1). */
uint64_t orc_siphash(int c_rounds, int d_rounds, uint64_t k0, uint64_t k1,
                     const uint8_t *msg, size_t len);
/* CuckooAddresser::hash_indices: out[i] = SipHash-1-3_0(LE64(seeds[i]) ||
 * key) % num_rows. num_rows = 0 is refused (the reference divides by it). */
int orc_cuckoo_indices(const uint8_t *key, size_t len, uint64_t num_rows,
                       const uint64_t seeds[3], uint64_t out[3]);

#ifdef __cplusplus
}
#endif

#endif
