/*
 * CPU oracle for the ChaCha8 DPF + masked-XOR PIR scan (test infrastructure
 * only; see oracle.h for what it restates and how each part is pinned).
 */
#include "oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- ChaCha --
 * chacha_prg.rs:92-144. RFC 8439 state layout: 4 constants, 8 key words,
 * counter, 3 nonce words; rounds/2 double rounds (column then diagonal);
 * feed-forward add of the initial state. */

static inline uint32_t rotl32(uint32_t x, int n) {
  return (x << n) | (x >> (32 - n));
}

#define QR(a, b, c, d)                                                         \
  do {                                                                         \
    a += b; d ^= a; d = rotl32(d, 16);                                         \
    c += d; b ^= c; b = rotl32(b, 12);                                         \
    a += b; d ^= a; d = rotl32(d, 8);                                          \
    c += d; b ^= c; b = rotl32(b, 7);                                          \
  } while (0)

void orc_chacha_block(const uint32_t key[8], uint32_t counter,
                      const uint32_t nonce[3], int rounds, uint32_t out[16]) {
  uint32_t init[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                       key[0],      key[1],      key[2],      key[3],
                       key[4],      key[5],      key[6],      key[7],
                       counter,     nonce[0],    nonce[1],    nonce[2]};
  uint32_t x[16];
  memcpy(x, init, sizeof x);
  for (int r = 0; r < rounds / 2; r++) {
    QR(x[0], x[4], x[8], x[12]);
    QR(x[1], x[5], x[9], x[13]);
    QR(x[2], x[6], x[10], x[14]);
    QR(x[3], x[7], x[11], x[15]);
    QR(x[0], x[5], x[10], x[15]);
    QR(x[1], x[6], x[11], x[12]);
    QR(x[2], x[7], x[8], x[13]);
    QR(x[3], x[4], x[9], x[14]);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + init[i];
}

/* chacha_prg.rs:153-178: key = seed||seed, nonce 0; block(counter 0) words
 * 0-3 are the left seed, 4-7 the right seed; block(counter 1) word 0 & 1 is
 * the left control bit, word 1 & 1 the right one. */
void orc_prg_expand(const uint32_t seed[4], int rounds, uint32_t left[4],
                    uint32_t right[4], uint8_t *left_t, uint8_t *right_t) {
  const uint32_t key[8] = {seed[0], seed[1], seed[2], seed[3],
                           seed[0], seed[1], seed[2], seed[3]};
  const uint32_t nonce[3] = {0, 0, 0};
  uint32_t b0[16], b1[16];
  orc_chacha_block(key, 0, nonce, rounds, b0);
  orc_chacha_block(key, 1, nonce, rounds, b1);
  for (int i = 0; i < 4; i++) {
    left[i] = b0[i];
    right[i] = b0[4 + i];
  }
  *left_t = (uint8_t)(b1[0] & 1u);
  *right_t = (uint8_t)(b1[1] & 1u);
}

#define PRG_ROUNDS 8

/* ------------------------------------------------------------ wire format --
 * dpf.rs:259-316: root_seed[16] | root_t | party | d | d x (seed_cw[16] |
 * t_cw_left | t_cw_right) | final_cw[16]; exact length 19 + 18d + 16. Seeds
 * are 4 little-endian u32 words (chacha_prg.rs:27-48). */

static inline uint32_t le32(const uint8_t *p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
         ((uint32_t)p[3] << 24);
}
static inline void put_le32(uint8_t *p, uint32_t v) {
  p[0] = (uint8_t)v;
  p[1] = (uint8_t)(v >> 8);
  p[2] = (uint8_t)(v >> 16);
  p[3] = (uint8_t)(v >> 24);
}
static void seed_from_bytes(const uint8_t *p, uint32_t s[4]) {
  for (int i = 0; i < 4; i++) s[i] = le32(p + 4 * i);
}
static void seed_to_bytes(const uint32_t s[4], uint8_t *p) {
  for (int i = 0; i < 4; i++) put_le32(p + 4 * i, s[i]);
}

size_t orc_key_wire_len(int domain_bits) {
  return 19 + 18 * (size_t)domain_bits + 16;
}

/* from_bytes accepts any u8 domain with the exact length (dpf.rs:277-290);
 * this evaluator additionally refuses d > 32, the ChaChaParams limit
 * (dpf.rs:65-78), since it cannot hold more correction words. */
int orc_key_from_bytes(const uint8_t *bytes, size_t len, orc_key *out) {
  if (len < 19 + 16) return ORC_ERR_KEY_FORMAT;
  int d = bytes[18];
  if (len != orc_key_wire_len(d)) return ORC_ERR_KEY_FORMAT;
  if (d > ORC_MAX_DOMAIN_BITS) return ORC_ERR_DOMAIN;
  memset(out, 0, sizeof *out);
  seed_from_bytes(bytes, out->root_seed);
  out->root_t = bytes[16];
  out->party = bytes[17];
  out->domain_bits = (uint8_t)d;
  size_t off = 19;
  for (int l = 0; l < d; l++) {
    seed_from_bytes(bytes + off, out->cw_seed[l]);
    out->cw_t_left[l] = bytes[off + 16];
    out->cw_t_right[l] = bytes[off + 17];
    off += 18;
  }
  seed_from_bytes(bytes + off, out->final_cw);
  return ORC_OK;
}

size_t orc_key_to_bytes(const orc_key *key, uint8_t *out) {
  seed_to_bytes(key->root_seed, out);
  out[16] = key->root_t;
  out[17] = key->party;
  out[18] = key->domain_bits;
  size_t off = 19;
  for (int l = 0; l < key->domain_bits; l++) {
    seed_to_bytes(key->cw_seed[l], out + off);
    out[off + 16] = key->cw_t_left[l];
    out[off + 17] = key->cw_t_right[l];
    off += 18;
  }
  seed_to_bytes(key->final_cw, out + off);
  return off + 16;
}

/* ------------------------------------------------------------------ eval --
 * One GGM step (dpf.rs:107-134): expand, take the child on `bit`, and if the
 * parent's control byte equals 1 XOR in the level's correction word. The
 * control value is a full byte: wire keys may carry any byte, and only the
 * comparison with 1 matters. */
static void ggm_step(const orc_key *key, int level, int bit, uint32_t seed[4],
                     uint8_t *t) {
  uint32_t l[4], r[4];
  uint8_t lt, rt;
  orc_prg_expand(seed, PRG_ROUNDS, l, r, &lt, &rt);
  uint32_t *child = bit ? r : l;
  uint8_t ct = bit ? rt : lt;
  if (*t == 1) {
    for (int i = 0; i < 4; i++) child[i] ^= key->cw_seed[level][i];
    ct ^= bit ? key->cw_t_right[level] : key->cw_t_left[level];
  }
  memcpy(seed, child, 16);
  *t = ct;
}

static void leaf_output(const orc_key *key, const uint32_t seed[4], uint8_t t,
                        uint32_t out[4]) {
  for (int i = 0; i < 4; i++)
    out[i] = (t == 1) ? (seed[i] ^ key->final_cw[i]) : seed[i];
}

/* dpf.rs:103-142: MSB-first path walk; leaf ^= final_cw iff t == 1. */
void orc_eval(const orc_key *key, uint64_t index, uint32_t out[4]) {
  uint32_t seed[4];
  memcpy(seed, key->root_seed, 16);
  uint8_t t = key->root_t;
  int d = key->domain_bits;
  for (int level = 0; level < d; level++) {
    int bit = (int)((index >> (d - 1 - level)) & 1u);
    ggm_step(key, level, bit, seed, &t);
  }
  leaf_output(key, seed, t, out);
}

/* The node on index's path at tree level `levels` (0 <= levels <= d): the
 * first `levels` steps of orc_eval (dpf.rs:107-134). */
void orc_path_node(const orc_key *key, uint64_t index, int levels,
                   uint32_t seed[4], uint8_t *t) {
  memcpy(seed, key->root_seed, 16);
  *t = key->root_t;
  int d = key->domain_bits;
  for (int level = 0; level < levels; level++) {
    int bit = (int)((index >> (d - 1 - level)) & 1u);
    ggm_step(key, level, bit, seed, t);
  }
}

/* dpf.rs:210-246: both children of every node; the left child fills the
 * first half of the output. Only element 0 of each leaf slice is written. */
static void expand_recursive(const orc_key *key, int level,
                             const uint32_t seed[4], uint8_t t, uint32_t *out,
                             uint64_t len) {
  if (level == key->domain_bits) {
    leaf_output(key, seed, t, out);
    return;
  }
  uint32_t l[4], r[4];
  uint8_t lt, rt;
  orc_prg_expand(seed, PRG_ROUNDS, l, r, &lt, &rt);
  if (t == 1) {
    for (int i = 0; i < 4; i++) {
      l[i] ^= key->cw_seed[level][i];
      r[i] ^= key->cw_seed[level][i];
    }
    lt ^= key->cw_t_left[level];
    rt ^= key->cw_t_right[level];
  }
  uint64_t half = len / 2;
  expand_recursive(key, level + 1, l, lt, out, half);
  expand_recursive(key, level + 1, r, rt, out + 4 * half, len - half);
}

/* dpf.rs:164-208: prefix walk to the aligned subtree root at level
 * d - log2(size) (saturating), then recursive expansion. */
int orc_eval_subtree(const orc_key *key, uint64_t start, uint64_t size,
                     uint32_t *out) {
  if (size == 0 || (size & (size - 1)) != 0) return ORC_ERR_SUBTREE;
  if (start % size != 0) return ORC_ERR_SUBTREE;
  int depth = __builtin_ctzll(size);
  int d = key->domain_bits;
  int root_level = d > depth ? d - depth : 0;
  uint32_t seed[4];
  memcpy(seed, key->root_seed, 16);
  uint8_t t = key->root_t;
  for (int level = 0; level < root_level; level++) {
    int bit = (int)((start >> (d - 1 - level)) & 1u);
    ggm_step(key, level, bit, seed, &t);
  }
  expand_recursive(key, root_level, seed, t, out, size);
  return ORC_OK;
}

/* dpf.rs:147-158 */
int orc_full_eval(const orc_key *key, uint32_t *out) {
  uint64_t n = 1ull << key->domain_bits;
  expand_recursive(key, 0, key->root_seed, key->root_t, out, n);
  return ORC_OK;
}

/* ---------------------------------------------------------------- keygen --
 * dpf.rs:322-432 with seeded root seeds (the reference draws them from
 * rand::random): seed0 / seed1 are splitmix64 words 0-1 / 2-3 of `seed`,
 * low half first. */
int orc_keygen(int domain_bits, uint64_t target, uint64_t seed, orc_key *k0,
               orc_key *k1) {
  if (domain_bits <= 0 || domain_bits > ORC_MAX_DOMAIN_BITS)
    return ORC_ERR_DOMAIN;
  if (target >> domain_bits) return ORC_ERR_INDEX;
  memset(k0, 0, sizeof *k0);
  memset(k1, 0, sizeof *k1);
  for (int i = 0; i < 2; i++) {
    uint64_t a = orc_splitmix64_word(seed, (uint64_t)i);
    uint64_t b = orc_splitmix64_word(seed, (uint64_t)(2 + i));
    k0->root_seed[2 * i] = (uint32_t)a;
    k0->root_seed[2 * i + 1] = (uint32_t)(a >> 32);
    k1->root_seed[2 * i] = (uint32_t)b;
    k1->root_seed[2 * i + 1] = (uint32_t)(b >> 32);
  }
  uint32_t s0[4], s1[4];
  memcpy(s0, k0->root_seed, 16);
  memcpy(s1, k1->root_seed, 16);
  uint8_t t0 = 0, t1 = 1;
  for (int level = 0; level < domain_bits; level++) {
    uint32_t l0[4], r0[4], l1[4], r1[4];
    uint8_t lt0, rt0, lt1, rt1;
    orc_prg_expand(s0, PRG_ROUNDS, l0, r0, &lt0, &rt0);
    orc_prg_expand(s1, PRG_ROUNDS, l1, r1, &lt1, &rt1);
    int bit = (int)((target >> (domain_bits - 1 - level)) & 1u);
    uint32_t *keep0 = bit ? r0 : l0, *lose0 = bit ? l0 : r0;
    uint32_t *keep1 = bit ? r1 : l1, *lose1 = bit ? l1 : r1;
    uint8_t keep_t0 = bit ? rt0 : lt0, lose_t0 = bit ? lt0 : rt0;
    uint8_t keep_t1 = bit ? rt1 : lt1, lose_t1 = bit ? lt1 : rt1;
    uint32_t cw[4];
    for (int i = 0; i < 4; i++) cw[i] = lose0[i] ^ lose1[i];
    uint8_t t_cw_lose = (uint8_t)(lose_t0 ^ lose_t1 ^ t0 ^ t1 ^ 1);
    uint8_t t_cw_keep = (uint8_t)(keep_t0 ^ keep_t1 ^ t0 ^ t1);
    memcpy(k0->cw_seed[level], cw, 16);
    k0->cw_t_left[level] = bit ? t_cw_lose : t_cw_keep;
    k0->cw_t_right[level] = bit ? t_cw_keep : t_cw_lose;
    uint8_t nt0 = keep_t0, nt1 = keep_t1;
    if (t0 == 1) {
      for (int i = 0; i < 4; i++) keep0[i] ^= cw[i];
      nt0 ^= t_cw_keep;
    }
    if (t1 == 1) {
      for (int i = 0; i < 4; i++) keep1[i] ^= cw[i];
      nt1 ^= t_cw_keep;
    }
    memcpy(s0, keep0, 16);
    memcpy(s1, keep1, 16);
    t0 = nt0;
    t1 = nt1;
  }
  for (int i = 0; i < 4; i++) k0->final_cw[i] = s0[i] ^ s1[i] ^ 0xFFFFFFFFu;
  k0->root_t = 0;
  k0->party = 0;
  k0->domain_bits = (uint8_t)domain_bits;
  memcpy(k1->cw_seed, k0->cw_seed, sizeof k0->cw_seed);
  memcpy(k1->cw_t_left, k0->cw_t_left, sizeof k0->cw_t_left);
  memcpy(k1->cw_t_right, k0->cw_t_right, sizeof k0->cw_t_right);
  memcpy(k1->final_cw, k0->final_cw, 16);
  k1->root_t = 1;
  k1->party = 1;
  k1->domain_bits = (uint8_t)domain_bits;
  return ORC_OK;
}

/* ----------------------------------------------------------------- scans --
 * kernel.rs:1243-1256 generalised from 4096-byte pages to W-byte rows: the
 * 16-byte mask ANDs every 16-byte lane of the row, and the result is XOR-ed
 * into the accumulator. */
static void xor_row_masked(uint8_t *acc, const uint8_t *row, uint32_t w,
                           const uint32_t mask[4]) {
  uint64_t m[2];
  m[0] = (uint64_t)mask[0] | ((uint64_t)mask[1] << 32);
  m[1] = (uint64_t)mask[2] | ((uint64_t)mask[3] << 32);
  for (uint32_t off = 0; off < w; off += 16) {
    uint64_t a[2], r[2];
    memcpy(a, acc + off, 16);
    memcpy(r, row + off, 16);
    a[0] ^= r[0] & m[0];
    a[1] ^= r[1] & m[1];
    memcpy(acc + off, a, 16);
  }
}

static int check_shape(const orc_key *key, uint64_t n_rows,
                       uint32_t row_bytes) {
  if (row_bytes == 0 || row_bytes % 16 != 0) return ORC_ERR_ROWS;
  if (key->domain_bits < 64 && (n_rows >> key->domain_bits) != 0 &&
      n_rows != (1ull << key->domain_bits))
    return ORC_ERR_ROWS;
  return ORC_OK;
}

/* kernel.rs:1374-1383: definition-level scan, one point eval per row. */
int orc_scan_brute(const orc_key *key, const uint8_t *rows, uint64_t n_rows,
                   uint32_t row_bytes, uint8_t *out) {
  int rc = check_shape(key, n_rows, row_bytes);
  if (rc) return rc;
  memset(out, 0, row_bytes);
  for (uint64_t i = 0; i < n_rows; i++) {
    uint32_t m[4];
    orc_eval(key, i, m);
    xor_row_masked(out, rows + i * row_bytes, row_bytes, m);
  }
  return ORC_OK;
}

/* kernel.rs:1276-1371: tiles of min(2048, 2^d) rows handed out dynamically
 * to worker threads (rayon's into_par_iter), eval_subtree per key per tile,
 * masked XOR per row, XOR reduction of the per-thread accumulators. The keys
 * form a flat list (the reference fixes 3); the tile is bounded by the
 * smallest key domain so every eval_subtree call stays aligned. */
typedef struct {
  const orc_key *keys;
  int n_keys;
  const uint8_t *rows;
  uint64_t n_rows;
  uint32_t row_bytes;
  uint64_t tile;
  uint64_t n_tiles;
  atomic_ullong next;
  uint8_t *acc; /* n_threads x n_keys x row_bytes */
} tiled_job;

typedef struct {
  tiled_job *job;
  int id;
} tiled_arg;

static void *tiled_worker(void *p) {
  tiled_arg *a = (tiled_arg *)p;
  tiled_job *j = a->job;
  uint32_t w = j->row_bytes;
  uint8_t *acc = j->acc + (size_t)a->id * j->n_keys * w;
  uint32_t *masks = (uint32_t *)malloc(j->tile * 16);
  for (;;) {
    uint64_t t = atomic_fetch_add(&j->next, 1);
    if (t >= j->n_tiles) break;
    uint64_t start = t * j->tile;
    uint64_t end = start + j->tile;
    if (end > j->n_rows) end = j->n_rows;
    for (int k = 0; k < j->n_keys; k++) {
      memset(masks, 0, j->tile * 16);
      orc_eval_subtree(&j->keys[k], start, j->tile, masks);
      for (uint64_t i = start; i < end; i++)
        xor_row_masked(acc + (size_t)k * w, j->rows + i * w, w,
                       masks + 4 * (i - start));
    }
  }
  free(masks);
  return NULL;
}

int orc_scan_tiled(const orc_key *keys, int n_keys, const uint8_t *rows,
                   uint64_t n_rows, uint32_t row_bytes, int n_threads,
                   uint8_t *out) {
  if (n_keys <= 0) return ORC_OK;
  int min_d = 64;
  for (int k = 0; k < n_keys; k++) {
    int rc = check_shape(&keys[k], n_rows, row_bytes);
    if (rc) return rc;
    if (keys[k].domain_bits < min_d) min_d = keys[k].domain_bits;
  }
  memset(out, 0, (size_t)n_keys * row_bytes);
  if (n_rows == 0) return ORC_OK;
  if (n_threads < 1) n_threads = 1;
  uint64_t tile = 2048;
  if (min_d < 11) tile = 1ull << min_d;
  tiled_job job;
  job.keys = keys;
  job.n_keys = n_keys;
  job.rows = rows;
  job.n_rows = n_rows;
  job.row_bytes = row_bytes;
  job.tile = tile;
  job.n_tiles = (n_rows + tile - 1) / tile;
  atomic_init(&job.next, 0);
  job.acc = (uint8_t *)calloc((size_t)n_threads * n_keys, row_bytes);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
  tiled_arg *args = (tiled_arg *)malloc(sizeof(tiled_arg) * n_threads);
  for (int i = 0; i < n_threads; i++) {
    args[i].job = &job;
    args[i].id = i;
    if (i > 0) pthread_create(&th[i], NULL, tiled_worker, &args[i]);
  }
  tiled_worker(&args[0]);
  for (int i = 1; i < n_threads; i++) pthread_join(th[i], NULL);
  size_t per = (size_t)n_keys * row_bytes;
  for (int i = 0; i < n_threads; i++)
    for (size_t b = 0; b < per; b++) out[b] ^= job.acc[(size_t)i * per + b];
  free(job.acc);
  free(th);
  free(args);
  return ORC_OK;
}

/* scan.rs:434-485 with one row per leaf: every delta entry contributes
 * diff & mask(row_idx), the 16 mask bytes repeated across the row. */
int orc_scan_delta(const orc_key *keys, int n_keys, const uint64_t *row_idx,
                   const uint8_t *diffs, uint64_t n_entries,
                   uint32_t row_bytes, uint8_t *out) {
  if (row_bytes == 0 || row_bytes % 16 != 0) return ORC_ERR_ROWS;
  memset(out, 0, (size_t)n_keys * row_bytes);
  for (uint64_t e = 0; e < n_entries; e++)
    for (int k = 0; k < n_keys; k++) {
      uint32_t m[4];
      orc_eval(&keys[k], row_idx[e], m);
      xor_row_masked(out + (size_t)k * row_bytes, diffs + e * row_bytes,
                     row_bytes, m);
    }
  return ORC_OK;
}

/* ------------------------------------------------------------ seeded data --
 * Counter-based data: 64-bit little-endian word j of the matrix (row-major,
 * rows of W bytes) is output j of a splitmix64 generator started at `seed`,
 * so any row can be regenerated independently. */
uint64_t orc_splitmix64_word(uint64_t seed, uint64_t j) {
  uint64_t z = seed + (j + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void orc_fill_rows(uint64_t seed, uint64_t first_row, uint64_t n_rows,
                   uint32_t row_bytes, uint8_t *out) {
  uint64_t words_per_row = row_bytes / 8;
  uint64_t j0 = first_row * words_per_row;
  uint64_t n = n_rows * words_per_row;
  for (uint64_t j = 0; j < n; j++) {
    uint64_t v = orc_splitmix64_word(seed, j0 + j);
    memcpy(out + 8 * j, &v, 8);
  }
}

/* chunked_matrix.rs:83-93: state = state * 6364136223846793005 + 1 before
 * each byte, byte = state >> 56, one continuous stream over the matrix.
 * orc_lcg_state_at jumps `steps` updates ahead by squaring the affine map. */
#define LCG_A 6364136223846793005ull

uint64_t orc_lcg_state_at(uint64_t seed, uint64_t steps) {
  uint64_t mul = 1, add = 0;       /* accumulated map x -> mul*x + add */
  uint64_t cur_mul = LCG_A, cur_add = 1;
  while (steps) {
    if (steps & 1) {
      mul = mul * cur_mul;
      add = add * cur_mul + cur_add;
    }
    cur_add = cur_add * cur_mul + cur_add;
    cur_mul = cur_mul * cur_mul;
    steps >>= 1;
  }
  return mul * seed + add;
}

void orc_lcg_fill(uint64_t seed, uint64_t first_byte, uint64_t n_bytes,
                  uint8_t *out) {
  uint64_t state = orc_lcg_state_at(seed, first_byte);
  for (uint64_t i = 0; i < n_bytes; i++) {
    state = state * LCG_A + 1;
    out[i] = (uint8_t)(state >> 56);
  }
}

/* ------------------------------------------------------------- SipHash -- */
/* The published SipHash-c-d: four 64-bit lanes initialised from the key and
 * the constants "somepseudorandomlygeneratedbytes"; each 8-byte little-endian
 * word m is absorbed as v3 ^= m, c SipRounds, v0 ^= m; the last word carries
 * the length mod 256 in its top byte; finalisation XORs 0xff into v2 and runs
 * d SipRounds. Rust's std DefaultHasher::new() is SipHash-1-3 with k0 = k1 = 0
 * (cuckoo.rs:33-37 hashes with it). */
static inline uint64_t rotl64(uint64_t x, int b) { return (x << b) | (x >> (64 - b)); }

static inline void sip_round(uint64_t v[4]) {
  v[0] += v[1]; v[1] = rotl64(v[1], 13); v[1] ^= v[0]; v[0] = rotl64(v[0], 32);
  v[2] += v[3]; v[3] = rotl64(v[3], 16); v[3] ^= v[2];
  v[0] += v[3]; v[3] = rotl64(v[3], 21); v[3] ^= v[0];
  v[2] += v[1]; v[1] = rotl64(v[1], 17); v[1] ^= v[2]; v[2] = rotl64(v[2], 32);
}

uint64_t orc_siphash(int c_rounds, int d_rounds, uint64_t k0, uint64_t k1,
                     const uint8_t *msg, size_t len) {
  uint64_t v[4] = {k0 ^ 0x736f6d6570736575ull, k1 ^ 0x646f72616e646f6dull,
                   k0 ^ 0x6c7967656e657261ull, k1 ^ 0x7465646279746573ull};
  size_t i = 0;
  for (; i + 8 <= len; i += 8) {
    uint64_t m = 0;
    for (int b = 0; b < 8; b++) m |= (uint64_t)msg[i + b] << (8 * b);
    v[3] ^= m;
    for (int r = 0; r < c_rounds; r++) sip_round(v);
    v[0] ^= m;
  }
  uint64_t last = (uint64_t)(len & 0xff) << 56;
  for (size_t b = 0; i + b < len; b++) last |= (uint64_t)msg[i + b] << (8 * b);
  v[3] ^= last;
  for (int r = 0; r < c_rounds; r++) sip_round(v);
  v[0] ^= last;
  v[2] ^= 0xff;
  for (int r = 0; r < d_rounds; r++) sip_round(v);
  return v[0] ^ v[1] ^ v[2] ^ v[3];
}

int orc_cuckoo_indices(const uint8_t *key, size_t len, uint64_t num_rows,
                       const uint64_t seeds[3], uint64_t out[3]) {
  if (num_rows == 0) return ORC_ERR_ROWS;
  uint8_t *buf = (uint8_t *)malloc(len + 8);
  if (!buf) return ORC_ERR_ROWS;
  for (int i = 0; i < 3; i++) {
    for (int b = 0; b < 8; b++) buf[b] = (uint8_t)(seeds[i] >> (8 * b));
    if (len) memcpy(buf + 8, key, len);
    out[i] = orc_siphash(1, 3, 0, 0, buf, len + 8) % num_rows;
  }
  free(buf);
  return ORC_OK;
}

