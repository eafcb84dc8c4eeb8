/*
 * Fast CPU forms of the oracle's scan. TEST INFRASTRUCTURE ONLY, like
 * oracle.c: the full-size checks and the timed CPU baseline use them; nothing
 * in the product library links, loads or calls them.
 *
 *  orc_eval_subtree_fast  ChaChaKey::eval_subtree (dpf.rs:164-246) expanded
 *                         breadth first: one ChaCha8 expansion
 *                         (chacha_prg.rs:92-178) runs 16 tree nodes side by
 *                         side in 64-byte vectors, one node per 32-bit lane.
 *  orc_scan_fast          eval_fused_3dpf_cpu's tiled scan (kernel.rs:1276-1371)
 *                         over rows in memory, with those masks.
 *  orc_scan_seeded        the same over the seeded matrix of orc_fill_rows,
 *                         regenerating rows tile by tile, so the 64 GiB and
 *                         75 GiB matrices are checked without a host copy.
 *  orc_scan_rows8         the reference CPU server scan's shape
 *                         (crates/morphogen-server/src/scan.rs:600-649: a fold
 *                         over row chunks on all threads, :721-895: eight rows
 *                         x 64 bytes per step, masks ANDed and the eight
 *                         products XOR-ed as a tree) with the 16-byte ChaCha
 *                         masks of the scalar orc_eval_subtree per chunk.
 * Pinned (tests/test_oracle.py): orc_eval_subtree_fast equals the scalar
 * orc_eval_subtree (itself pinned by OpenSSL's ChaCha20 and the DPF
 * invariants) on random and fuzzed keys; every scan equals orc_scan_tiled /
 * orc_scan_brute on the same keys and rows; orc_scan_seeded equals
 * orc_scan_tiled over orc_fill_rows, whole and for row ranges.
 *
 * The vector code is written with GCC vector extensions and built for
 * AVX-512, AVX2 and plain x86-64 (target_clones), picked at load time, so the
 * library built in one place runs on any x86-64 host.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef uint32_t v16u __attribute__((vector_size(64)));
typedef uint32_t v8u __attribute__((vector_size(32)));
typedef uint32_t v4u __attribute__((vector_size(16)));
typedef uint64_t v8q __attribute__((vector_size(64)));

#define LANES 16
#define CLONES __attribute__((target_clones("avx512f", "avx2", "default")))
#define INLINE static inline __attribute__((always_inline))

/* ------------------------------------------------- 16-lane ChaCha8 PRG -- */

INLINE v16u rotv(v16u x, int n) { return (x << n) | (x >> (32 - n)); }

#define QRV(a, b, c, d)                                                        \
  do {                                                                         \
    a += b; d ^= a; d = rotv(d, 16);                                           \
    c += d; b ^= c; b = rotv(b, 12);                                           \
    a += b; d ^= a; d = rotv(d, 8);                                            \
    c += d; b ^= c; b = rotv(b, 7);                                            \
  } while (0)

/* One ChaCha8 block per lane, key = seed||seed, nonce 0 (chacha_prg.rs:
 * 153-178 builds the key that way); out = state after the feed-forward. */
INLINE void chacha8_x16(const v16u s[4], uint32_t counter, v16u out[16]) {
  const v16u zero = {0};
  v16u init[16];
  init[0] = zero + 0x61707865u;
  init[1] = zero + 0x3320646eu;
  init[2] = zero + 0x79622d32u;
  init[3] = zero + 0x6b206574u;
  for (int i = 0; i < 4; i++) {
    init[4 + i] = s[i];
    init[8 + i] = s[i];
  }
  init[12] = zero + counter;
  init[13] = zero;
  init[14] = zero;
  init[15] = zero;
  v16u x[16];
  for (int i = 0; i < 16; i++) x[i] = init[i];
  for (int r = 0; r < 4; r++) {
    QRV(x[0], x[4], x[8], x[12]);
    QRV(x[1], x[5], x[9], x[13]);
    QRV(x[2], x[6], x[10], x[14]);
    QRV(x[3], x[7], x[11], x[15]);
    QRV(x[0], x[5], x[10], x[15]);
    QRV(x[1], x[6], x[11], x[12]);
    QRV(x[2], x[7], x[8], x[13]);
    QRV(x[3], x[4], x[9], x[14]);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + init[i];
}

/* Nodes as structure of arrays: seed words w[0..3] and the control byte t
 * (kept as a u32 holding the byte, any value a wire key can carry). */
typedef struct {
  uint32_t *w[4];
  uint32_t *t;
} soa;

static void soa_at(soa *s, uint32_t *base, uint32_t cap) {
  for (int j = 0; j < 4; j++) s->w[j] = base + (size_t)j * cap;
  s->t = base + (size_t)4 * cap;
}

/* One tree level (dpf.rs:210-246 for every node of the level at once): node
 * i of src (m nodes at tree level `level`) expands into its left child at
 * dst[i] and its right child at dst[m + i]; children take the level's
 * correction word when the parent's control byte is 1. With this layout the
 * node at position p after L levels is the one at offset bitrev_L(p). */
CLONES
static void expand_level(const orc_key *key, int level, const soa *src, const soa *dst,
                         uint32_t m) {
  const v16u zero = {0};
  v16u cw[4];
  for (int j = 0; j < 4; j++) cw[j] = zero + key->cw_seed[level][j];
  const v16u ctl = zero + key->cw_t_left[level];
  const v16u ctr = zero + key->cw_t_right[level];
  for (uint32_t i = 0; i < m; i += LANES) {
    const uint32_t n = m - i < LANES ? m - i : LANES;
    v16u s[4], tv;
    if (n == LANES) {
      for (int j = 0; j < 4; j++) memcpy(&s[j], src->w[j] + i, 64);
      memcpy(&tv, src->t + i, 64);
    } else {
      for (int j = 0; j < 4; j++) {
        s[j] = zero;
        memcpy(&s[j], src->w[j] + i, 4 * n);
      }
      tv = zero;
      memcpy(&tv, src->t + i, 4 * n);
    }
    v16u b0[16], b1[16];
    chacha8_x16(s, 0, b0);
    chacha8_x16(s, 1, b1);
    const v16u on = (v16u)(tv == 1);
    v16u l[4], r[4];
    for (int j = 0; j < 4; j++) {
      l[j] = b0[j] ^ (cw[j] & on);
      r[j] = b0[4 + j] ^ (cw[j] & on);
    }
    const v16u lt = (b1[0] & 1u) ^ (ctl & on);
    const v16u rt = (b1[1] & 1u) ^ (ctr & on);
    for (int j = 0; j < 4; j++) {
      memcpy(dst->w[j] + i, &l[j], 4 * n);
      memcpy(dst->w[j] + m + i, &r[j], 4 * n);
    }
    memcpy(dst->t + i, &lt, 4 * n);
    memcpy(dst->t + m + i, &rt, 4 * n);
  }
}

static uint32_t bitrev(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

/* Scratch of subtree_masks for 2^levels leaves: two node arrays. */
static size_t subtree_scratch_words(int levels) { return (size_t)10 << levels; }

/* Leaf masks of the aligned 2^levels-leaf subtree whose root (seed, t) sits
 * at tree level d - levels, in natural order: masks[4x + j] is word j of the
 * leaf at offset x (leaf output of dpf.rs:136-141: seed ^ final_cw iff
 * t == 1). `rev` maps positions to offsets. */
static void subtree_masks(const orc_key *key, const uint32_t seed[4], uint8_t t, int levels,
                          uint32_t *scratch, const uint32_t *rev, uint32_t *masks) {
  const uint32_t cap = 1u << levels;
  soa a, b;
  soa_at(&a, scratch, cap);
  soa_at(&b, scratch + (size_t)5 * cap, cap);
  for (int j = 0; j < 4; j++) a.w[j][0] = seed[j];
  a.t[0] = t;
  const int level0 = key->domain_bits - levels;
  uint32_t m = 1;
  for (int l = 0; l < levels; l++) {
    expand_level(key, level0 + l, &a, &b, m);
    soa tmp = a;
    a = b;
    b = tmp;
    m <<= 1;
  }
  for (uint32_t p = 0; p < cap; p++) {
    const uint32_t x = rev[p];
    const int on = a.t[p] == 1;
    for (int j = 0; j < 4; j++) masks[4 * (size_t)x + j] = a.w[j][p] ^ (on ? key->final_cw[j] : 0u);
  }
}

static uint32_t *rev_table(int levels) {
  uint32_t *rev = (uint32_t *)malloc(sizeof(uint32_t) << levels);
  if (rev)
    for (uint32_t p = 0; p < (1u << levels); p++) rev[p] = bitrev(p, levels);
  return rev;
}

int orc_eval_subtree_fast(const orc_key *key, uint64_t start, uint64_t size, uint32_t *out) {
  if (size == 0 || (size & (size - 1)) != 0) return ORC_ERR_SUBTREE;
  if (start % size != 0) return ORC_ERR_SUBTREE;
  const int levels = __builtin_ctzll(size);
  if (levels > key->domain_bits || levels > 20) return ORC_ERR_SUBTREE;
  uint32_t seed[4];
  uint8_t t;
  orc_path_node(key, start, key->domain_bits - levels, seed, &t);
  uint32_t *scratch = (uint32_t *)malloc(sizeof(uint32_t) * subtree_scratch_words(levels));
  uint32_t *rev = rev_table(levels);
  if (!scratch || !rev) {
    free(scratch);
    free(rev);
    return ORC_ERR_SUBTREE;
  }
  subtree_masks(key, seed, t, levels, scratch, rev, out);
  free(scratch);
  free(rev);
  return ORC_OK;
}

/* ------------------------------------------------------ masked XOR fold -- */

/* acc ^= rows[r] & tile16(masks[r]) for n rows of w bytes (kernel.rs:1243-1256). */
CLONES
static void fold_rows(uint8_t *acc, const uint8_t *rows, uint32_t w, uint32_t n,
                      const uint32_t *masks) {
  for (uint32_t r = 0; r < n; r++) {
    const uint32_t *m = masks + 4 * (size_t)r;
    const uint8_t *row = rows + (size_t)r * w;
    uint32_t off = 0;
    if (w >= 64) {
      const v16u mv = {m[0], m[1], m[2], m[3], m[0], m[1], m[2], m[3],
                       m[0], m[1], m[2], m[3], m[0], m[1], m[2], m[3]};
      for (; off + 64 <= w; off += 64) {
        v16u a, x;
        memcpy(&a, acc + off, 64);
        memcpy(&x, row + off, 64);
        a ^= x & mv;
        memcpy(acc + off, &a, 64);
      }
    }
    const v4u m4 = {m[0], m[1], m[2], m[3]};
    for (; off < w; off += 16) {
      v4u a, x;
      memcpy(&a, acc + off, 16);
      memcpy(&x, row + off, 16);
      a ^= x & m4;
      memcpy(acc + off, &a, 16);
    }
  }
}

/* Words [j0, j0 + n) of the seeded matrix (orc_splitmix64_word). */
CLONES
static void fill_words(uint64_t seed, uint64_t j0, uint64_t n, uint64_t *out) {
  for (uint64_t j = 0; j < n; j++) {
    uint64_t z = seed + (j0 + j + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    out[j] = z ^ (z >> 31);
  }
}

/* ------------------------------------------------------ tiled fast scan -- */

/* Consecutive tiles go to a thread in runs, and each key keeps the path of
 * its last tile root, so a new tile re-walks only the levels below the first
 * index bit that changed (about two node expansions per tile). */
typedef struct {
  uint32_t seed[ORC_MAX_DOMAIN_BITS + 1][4];
  uint8_t t[ORC_MAX_DOMAIN_BITS + 1];
  uint64_t prefix;
  int valid;
} path_cache;

static void tile_root(const orc_key *key, path_cache *pc, uint64_t tile_start, int levels,
                      uint32_t seed[4], uint8_t *t) {
  const int d = key->domain_bits;
  const int top = d - levels;  /* tree level of the tile root */
  const uint64_t prefix = top ? tile_start >> levels : 0;
  int from = 0;
  if (pc->valid) {
    if (pc->prefix == prefix) {
      from = top;
    } else {
      const uint64_t diff = pc->prefix ^ prefix;
      from = top - (64 - __builtin_clzll(diff));
    }
  } else {
    memcpy(pc->seed[0], key->root_seed, 16);
    pc->t[0] = key->root_t;
  }
  for (int l = from; l < top; l++) {
    uint32_t lft[4], rgt[4];
    uint8_t lt, rt;
    orc_prg_expand(pc->seed[l], 8, lft, rgt, &lt, &rt);
    const int bit = (int)((prefix >> (top - 1 - l)) & 1u);
    uint32_t *c = bit ? rgt : lft;
    uint8_t ct = bit ? rt : lt;
    if (pc->t[l] == 1) {
      for (int j = 0; j < 4; j++) c[j] ^= key->cw_seed[l][j];
      ct ^= bit ? key->cw_t_right[l] : key->cw_t_left[l];
    }
    memcpy(pc->seed[l + 1], c, 16);
    pc->t[l + 1] = ct;
  }
  pc->prefix = prefix;
  pc->valid = 1;
  memcpy(seed, pc->seed[top], 16);
  *t = pc->t[top];
}

enum { RUN_TILES = 16, FOLD_ROWS = 64 };

typedef struct {
  const orc_key *keys;
  int n_keys;
  const uint8_t *rows; /* NULL: seeded rows */
  uint64_t seed;
  uint64_t row_begin, row_end;
  uint32_t w;
  int levels;
  uint64_t tile0, n_tiles;
  const uint32_t *rev;
  atomic_ullong next;
  uint8_t *acc; /* n_threads x n_keys x w */
  int failed;
} fast_job;

typedef struct {
  fast_job *job;
  int id;
} fast_arg;

static void *fast_worker(void *p) {
  fast_arg *a = (fast_arg *)p;
  fast_job *j = a->job;
  const uint32_t w = j->w;
  const uint64_t tile_rows = 1ull << j->levels;
  uint8_t *acc = j->acc + (size_t)a->id * j->n_keys * w;
  uint32_t *scratch = (uint32_t *)malloc(sizeof(uint32_t) * subtree_scratch_words(j->levels));
  uint32_t *masks = (uint32_t *)malloc((size_t)j->n_keys * tile_rows * 16);
  path_cache *pc = (path_cache *)calloc((size_t)j->n_keys, sizeof(path_cache));
  uint64_t *rowbuf = j->rows ? NULL : (uint64_t *)malloc((size_t)FOLD_ROWS * w);
  if (!scratch || !masks || !pc || (!j->rows && !rowbuf)) {
    j->failed = 1;
    goto out;
  }
  for (;;) {
    const uint64_t r0 = atomic_fetch_add(&j->next, RUN_TILES);
    if (r0 >= j->n_tiles) break;
    const uint64_t r1 = r0 + RUN_TILES < j->n_tiles ? r0 + RUN_TILES : j->n_tiles;
    for (int k = 0; k < j->n_keys; k++) pc[k].valid = 0;
    for (uint64_t ti = r0; ti < r1; ti++) {
      const uint64_t tile_start = (j->tile0 + ti) * tile_rows;
      uint64_t lo = tile_start > j->row_begin ? tile_start : j->row_begin;
      uint64_t hi = tile_start + tile_rows < j->row_end ? tile_start + tile_rows : j->row_end;
      for (int k = 0; k < j->n_keys; k++) {
        uint32_t seed[4];
        uint8_t t;
        tile_root(&j->keys[k], &pc[k], tile_start, j->levels, seed, &t);
        subtree_masks(&j->keys[k], seed, t, j->levels, scratch, j->rev,
                      masks + (size_t)k * tile_rows * 4);
      }
      for (uint64_t x = lo; x < hi; x += FOLD_ROWS) {
        const uint32_t n = (uint32_t)(hi - x < FOLD_ROWS ? hi - x : FOLD_ROWS);
        const uint8_t *rows;
        if (j->rows) {
          rows = j->rows + (x - j->row_begin) * w;
        } else {
          fill_words(j->seed, x * (w / 8), (uint64_t)n * (w / 8), rowbuf);
          rows = (const uint8_t *)rowbuf;
        }
        for (int k = 0; k < j->n_keys; k++)
          fold_rows(acc + (size_t)k * w, rows, w, n,
                    masks + ((size_t)k * tile_rows + (x - tile_start)) * 4);
      }
    }
  }
out:
  free(scratch);
  free(masks);
  free(pc);
  free(rowbuf);
  return NULL;
}

/* Rows [row_begin, row_end) of an n_rows-row matrix: from `rows` (holding
 * exactly that range) or, with rows NULL, regenerated from `seed`. */
static int scan_fast_range(const orc_key *keys, int n_keys, const uint8_t *rows, uint64_t seed,
                           uint64_t n_rows, uint64_t row_begin, uint64_t row_end, uint32_t w,
                           int n_threads, uint8_t *out) {
  if (n_keys <= 0) return ORC_OK;
  if (w == 0 || w % 16 != 0) return ORC_ERR_ROWS;
  if (row_begin > row_end || row_end > n_rows) return ORC_ERR_ROWS;
  int min_d = 64;
  for (int k = 0; k < n_keys; k++) {
    const int d = keys[k].domain_bits;
    if (d < 64 && (n_rows >> d) != 0 && n_rows != (1ull << d)) return ORC_ERR_ROWS;
    if (d < min_d) min_d = d;
  }
  for (int k = 0; k < n_keys; k++)
    if (keys[k].domain_bits != min_d) return ORC_ERR_DOMAIN; /* one tile depth for all keys */
  memset(out, 0, (size_t)n_keys * w);
  if (row_end == row_begin) return ORC_OK;
  if (n_threads < 1) n_threads = 1;
  fast_job job;
  memset(&job, 0, sizeof job);
  job.keys = keys;
  job.n_keys = n_keys;
  job.rows = rows;
  job.seed = seed;
  job.row_begin = row_begin;
  job.row_end = row_end;
  job.w = w;
  job.levels = min_d < 12 ? min_d : 12;
  const uint64_t tile_rows = 1ull << job.levels;
  job.tile0 = row_begin / tile_rows;
  job.n_tiles = (row_end + tile_rows - 1) / tile_rows - job.tile0;
  uint32_t *rev = rev_table(job.levels);
  job.rev = rev;
  atomic_init(&job.next, 0);
  job.acc = (uint8_t *)calloc((size_t)n_threads * n_keys, w);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
  fast_arg *args = (fast_arg *)malloc(sizeof(fast_arg) * n_threads);
  if (!rev || !job.acc || !th || !args) {
    free(rev);
    free(job.acc);
    free(th);
    free(args);
    return ORC_ERR_ROWS;
  }
  for (int i = 0; i < n_threads; i++) {
    args[i].job = &job;
    args[i].id = i;
    if (i > 0) pthread_create(&th[i], NULL, fast_worker, &args[i]);
  }
  fast_worker(&args[0]);
  for (int i = 1; i < n_threads; i++) pthread_join(th[i], NULL);
  const size_t per = (size_t)n_keys * w;
  for (int i = 0; i < n_threads; i++)
    for (size_t b = 0; b < per; b++) out[b] ^= job.acc[(size_t)i * per + b];
  const int failed = job.failed;
  free(rev);
  free(job.acc);
  free(th);
  free(args);
  return failed ? ORC_ERR_ROWS : ORC_OK;
}

int orc_scan_fast(const orc_key *keys, int n_keys, const uint8_t *rows, uint64_t n_rows,
                  uint32_t row_bytes, int n_threads, uint8_t *out) {
  return scan_fast_range(keys, n_keys, rows, 0, n_rows, 0, n_rows, row_bytes, n_threads, out);
}

int orc_scan_seeded(const orc_key *keys, int n_keys, uint64_t seed, uint64_t n_rows,
                    uint64_t row_begin, uint64_t row_end, uint32_t row_bytes, int n_threads,
                    uint8_t *out) {
  if (row_bytes % 8 != 0) return ORC_ERR_ROWS;
  return scan_fast_range(keys, n_keys, NULL, seed, n_rows, row_begin, row_end, row_bytes,
                         n_threads, out);
}

/* ------------------------------------------- reference-shaped rows8 scan -- */

enum { ROWS8_CHUNK = 2048 };

/* Eight rows per step (scan.rs:721-895): for each 64-byte column block, load
 * the eight rows' blocks once, AND each with its row's mask per key, XOR the
 * eight products as a tree and fold them into the key's accumulator. Rows
 * narrower than a 64-byte block, and the tail of wider ones, go through the
 * same steps with 32- or 16-byte vectors. */
CLONES
static void rows8_fold(uint8_t *acc, int n_keys, const uint8_t *rows, uint32_t w, uint64_t n,
                       const uint32_t *masks, uint64_t mask_stride) {
  uint64_t r = 0;
  for (; r + 8 <= n; r += 8) {
    const uint8_t *rp = rows + r * w;
    uint32_t off = 0;
    for (; off + 64 <= w; off += 64) {
      v16u d[8];
      for (int j = 0; j < 8; j++) memcpy(&d[j], rp + (size_t)j * w + off, 64);
      for (int k = 0; k < n_keys; k++) {
        v16u mm[8];
        for (int j = 0; j < 8; j++) {
          const uint32_t *m = masks + (k * mask_stride + r + j) * 4;
          mm[j] = (v16u){m[0], m[1], m[2], m[3], m[0], m[1], m[2], m[3],
                         m[0], m[1], m[2], m[3], m[0], m[1], m[2], m[3]};
        }
        const v16u x = (((d[0] & mm[0]) ^ (d[1] & mm[1])) ^ ((d[2] & mm[2]) ^ (d[3] & mm[3]))) ^
                       (((d[4] & mm[4]) ^ (d[5] & mm[5])) ^ ((d[6] & mm[6]) ^ (d[7] & mm[7])));
        v16u a;
        memcpy(&a, acc + (size_t)k * w + off, 64);
        a ^= x;
        memcpy(acc + (size_t)k * w + off, &a, 64);
      }
    }
    for (; off + 32 <= w; off += 32) {
      v8u d[8];
      for (int j = 0; j < 8; j++) memcpy(&d[j], rp + (size_t)j * w + off, 32);
      for (int k = 0; k < n_keys; k++) {
        v8u mm[8];
        for (int j = 0; j < 8; j++) {
          const uint32_t *m = masks + (k * mask_stride + r + j) * 4;
          mm[j] = (v8u){m[0], m[1], m[2], m[3], m[0], m[1], m[2], m[3]};
        }
        const v8u x = (((d[0] & mm[0]) ^ (d[1] & mm[1])) ^ ((d[2] & mm[2]) ^ (d[3] & mm[3]))) ^
                      (((d[4] & mm[4]) ^ (d[5] & mm[5])) ^ ((d[6] & mm[6]) ^ (d[7] & mm[7])));
        v8u a;
        memcpy(&a, acc + (size_t)k * w + off, 32);
        a ^= x;
        memcpy(acc + (size_t)k * w + off, &a, 32);
      }
    }
    for (; off < w; off += 16) {
      for (int k = 0; k < n_keys; k++) {
        v4u x = {0, 0, 0, 0};
        for (int j = 0; j < 8; j++) {
          const uint32_t *m = masks + (k * mask_stride + r + j) * 4;
          v4u dj;
          memcpy(&dj, rp + (size_t)j * w + off, 16);
          x ^= dj & (v4u){m[0], m[1], m[2], m[3]};
        }
        v4u a;
        memcpy(&a, acc + (size_t)k * w + off, 16);
        a ^= x;
        memcpy(acc + (size_t)k * w + off, &a, 16);
      }
    }
  }
  for (; r < n; r++) /* remaining rows one at a time */
    for (int k = 0; k < n_keys; k++)
      fold_rows(acc + (size_t)k * w, rows + r * w, w, 1, masks + (k * mask_stride + r) * 4);
}

typedef struct {
  const orc_key *keys;
  int n_keys;
  const uint8_t *rows;
  uint64_t n_rows;
  uint32_t w;
  uint64_t chunk;
  uint64_t n_chunks;
  atomic_ullong next;
  uint8_t *acc;
  int failed;
} rows8_job;

typedef struct {
  rows8_job *job;
  int id;
} rows8_arg;

static void *rows8_worker(void *p) {
  rows8_arg *a = (rows8_arg *)p;
  rows8_job *j = a->job;
  uint8_t *acc = j->acc + (size_t)a->id * j->n_keys * j->w;
  uint32_t *masks = (uint32_t *)malloc((size_t)j->n_keys * j->chunk * 16);
  if (!masks) {
    j->failed = 1;
    return NULL;
  }
  for (;;) {
    const uint64_t c = atomic_fetch_add(&j->next, 1);
    if (c >= j->n_chunks) break;
    const uint64_t start = c * j->chunk;
    const uint64_t end = start + j->chunk < j->n_rows ? start + j->chunk : j->n_rows;
    for (int k = 0; k < j->n_keys; k++)
      orc_eval_subtree(&j->keys[k], start, j->chunk, masks + (size_t)k * j->chunk * 4);
    rows8_fold(acc, j->n_keys, j->rows + start * j->w, j->w, end - start, masks, j->chunk);
  }
  free(masks);
  return NULL;
}

int orc_scan_rows8(const orc_key *keys, int n_keys, const uint8_t *rows, uint64_t n_rows,
                   uint32_t row_bytes, int n_threads, uint8_t *out) {
  if (n_keys <= 0) return ORC_OK;
  if (row_bytes == 0 || row_bytes % 16 != 0) return ORC_ERR_ROWS;
  int min_d = 64;
  for (int k = 0; k < n_keys; k++) {
    const int d = keys[k].domain_bits;
    if (d < 64 && (n_rows >> d) != 0 && n_rows != (1ull << d)) return ORC_ERR_ROWS;
    if (d < min_d) min_d = d;
  }
  memset(out, 0, (size_t)n_keys * row_bytes);
  if (n_rows == 0) return ORC_OK;
  if (n_threads < 1) n_threads = 1;
  rows8_job job;
  memset(&job, 0, sizeof job);
  job.keys = keys;
  job.n_keys = n_keys;
  job.rows = rows;
  job.n_rows = n_rows;
  job.w = row_bytes;
  job.chunk = min_d < 11 ? (1ull << min_d) : ROWS8_CHUNK;
  job.n_chunks = (n_rows + job.chunk - 1) / job.chunk;
  atomic_init(&job.next, 0);
  job.acc = (uint8_t *)calloc((size_t)n_threads * n_keys, row_bytes);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
  rows8_arg *args = (rows8_arg *)malloc(sizeof(rows8_arg) * n_threads);
  if (!job.acc || !th || !args) {
    free(job.acc);
    free(th);
    free(args);
    return ORC_ERR_ROWS;
  }
  for (int i = 0; i < n_threads; i++) {
    args[i].job = &job;
    args[i].id = i;
    if (i > 0) pthread_create(&th[i], NULL, rows8_worker, &args[i]);
  }
  rows8_worker(&args[0]);
  for (int i = 1; i < n_threads; i++) pthread_join(th[i], NULL);
  const size_t per = (size_t)n_keys * row_bytes;
  for (int i = 0; i < n_threads; i++)
    for (size_t b = 0; b < per; b++) out[b] ^= job.acc[(size_t)i * per + b];
  const int failed = job.failed;
  free(job.acc);
  free(th);
  free(args);
  return failed ? ORC_ERR_ROWS : ORC_OK;
}

/* Which vector width the clones above run with on this host. */
const char *orc_simd_isa(void) {
  __builtin_cpu_init();
  if (__builtin_cpu_supports("avx512f")) return "avx512f";
  if (__builtin_cpu_supports("avx2")) return "avx2";
  return "x86-64";
}
