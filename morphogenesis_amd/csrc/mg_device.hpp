// Device-side ChaCha8 PRG, GGM step and key layout for gfx950.
//
// Semantics follow crates/morphogen-gpu-dpf/src/chacha_prg.rs:92-178 (block,
// expand) and src/dpf.rs:103-142 (one GGM step with correction words); the
// code is written for the CDNA4 VALU: every word lives in a VGPR, rotates are
// single v_alignbit_b32 ops and `x ^ (y & m)` folds into one v_bitop3_b32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mg_key.hpp"

namespace mg {

struct Node {
  uint32_t s0, s1, s2, s3;
  uint32_t t;  // control byte; only `t == 1` triggers a correction
};

#define MG_ROTL(x, n) (((x) << (n)) | ((x) >> (32 - (n))))
#define MG_QR(a, b, c, d)                                                      \
  do {                                                                         \
    a += b; d ^= a; d = MG_ROTL(d, 16);                                        \
    c += d; b ^= c; b = MG_ROTL(b, 12);                                        \
    a += b; d ^= a; d = MG_ROTL(d, 8);                                         \
    c += d; b ^= c; b = MG_ROTL(b, 7);                                         \
  } while (0)

constexpr uint32_t kC0 = 0x61707865u, kC1 = 0x3320646eu, kC2 = 0x79622d32u,
                   kC3 = 0x6b206574u;

// PRG expand (chacha_prg.rs:153-178): key = seed||seed, nonce 0. Block 0
// gives the two child seeds (words 0-3, 4-7); block 1 only contributes the
// low bits of words 0 and 1. Both blocks are written out in full; the
// compiler drops block 1's unused words and shares its three counter-free
// first-round column quarter-rounds with block 0.
__device__ __forceinline__ void prg_expand(uint32_t s0, uint32_t s1,
                                           uint32_t s2, uint32_t s3,
                                           uint32_t l[4], uint32_t r[4],
                                           uint32_t& lt, uint32_t& rt) {
  uint32_t x0 = kC0, x1 = kC1, x2 = kC2, x3 = kC3;
  uint32_t x4 = s0, x5 = s1, x6 = s2, x7 = s3;
  uint32_t x8 = s0, x9 = s1, x10 = s2, x11 = s3;
  uint32_t x12 = 0, x13 = 0, x14 = 0, x15 = 0;
  uint32_t y0 = kC0, y1 = kC1, y2 = kC2, y3 = kC3;
  uint32_t y4 = s0, y5 = s1, y6 = s2, y7 = s3;
  uint32_t y8 = s0, y9 = s1, y10 = s2, y11 = s3;
  uint32_t y12 = 1, y13 = 0, y14 = 0, y15 = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    MG_QR(x0, x4, x8, x12);
    MG_QR(x1, x5, x9, x13);
    MG_QR(x2, x6, x10, x14);
    MG_QR(x3, x7, x11, x15);
    MG_QR(x0, x5, x10, x15);
    MG_QR(x1, x6, x11, x12);
    MG_QR(x2, x7, x8, x13);
    MG_QR(x3, x4, x9, x14);
    MG_QR(y0, y4, y8, y12);
    MG_QR(y1, y5, y9, y13);
    MG_QR(y2, y6, y10, y14);
    MG_QR(y3, y7, y11, y15);
    MG_QR(y0, y5, y10, y15);
    MG_QR(y1, y6, y11, y12);
    MG_QR(y2, y7, y8, y13);
    MG_QR(y3, y4, y9, y14);
  }
  l[0] = x0 + kC0; l[1] = x1 + kC1; l[2] = x2 + kC2; l[3] = x3 + kC3;
  r[0] = x4 + s0; r[1] = x5 + s1; r[2] = x6 + s2; r[3] = x7 + s3;
  lt = (y0 + kC0) & 1u;
  rt = (y1 + kC1) & 1u;
}

// All-ones if t == 1, else zero, kept opaque to the optimiser so that
// `x ^ (uniform & m)` stays one v_bitop3_b32 with the uniform operand in an
// SGPR, instead of being rewritten as a select of the uniform value (which
// costs a v_mov plus a v_cndmask per word).
__device__ __forceinline__ uint32_t t_mask(uint32_t t) {
  uint32_t m = (t == 1u) ? 0xFFFFFFFFu : 0u;
  asm("" : "+v"(m));
  return m;
}

// Both children of `n` at tree level `level`, correction word applied when
// n.t == 1 (dpf.rs:221-239). `k` is wave-uniform, so the correction word is
// read with scalar loads.
__device__ __forceinline__ void expand_node(const DevKey* __restrict__ k,
                                            int level, const Node& n,
                                            Node& left, Node& right) {
  uint32_t l[4], r[4], lt, rt;
  prg_expand(n.s0, n.s1, n.s2, n.s3, l, r, lt, rt);
  const uint32_t m = t_mask(n.t);
  const uint32_t c0 = k->cw_seed[level][0], c1 = k->cw_seed[level][1];
  const uint32_t c2 = k->cw_seed[level][2], c3 = k->cw_seed[level][3];
  const uint32_t ct = k->cw_t[level];
  left.s0 = l[0] ^ (c0 & m); left.s1 = l[1] ^ (c1 & m);
  left.s2 = l[2] ^ (c2 & m); left.s3 = l[3] ^ (c3 & m);
  right.s0 = r[0] ^ (c0 & m); right.s1 = r[1] ^ (c1 & m);
  right.s2 = r[2] ^ (c2 & m); right.s3 = r[3] ^ (c3 & m);
  left.t = lt ^ (ct & 0xFFu & m);
  right.t = rt ^ ((ct >> 8) & 0xFFu & m);
}

// One step down the tree towards child `bit` (dpf.rs:107-134).
__device__ __forceinline__ Node ggm_step(const DevKey* __restrict__ k,
                                         int level, const Node& n,
                                         uint32_t bit) {
  Node a, b;
  expand_node(k, level, n, a, b);
  return bit ? b : a;
}

// Leaf output (dpf.rs:136-141): seed ^ final_cw iff t == 1.
__device__ __forceinline__ uint4 leaf_mask(const DevKey* __restrict__ k,
                                           const Node& n) {
  const uint32_t m = t_mask(n.t);
  return make_uint4(n.s0 ^ (k->final_cw[0] & m), n.s1 ^ (k->final_cw[1] & m),
                    n.s2 ^ (k->final_cw[2] & m), n.s3 ^ (k->final_cw[3] & m));
}

__device__ __forceinline__ Node key_root(const DevKey* __restrict__ k) {
  Node n;
  n.s0 = k->root_seed[0]; n.s1 = k->root_seed[1];
  n.s2 = k->root_seed[2]; n.s3 = k->root_seed[3];
  n.t = k->root_t;
  return n;
}

// Full path walk to leaf `x` (dpf.rs:103-142).
__device__ __forceinline__ uint4 point_eval(const DevKey* __restrict__ k,
                                            uint64_t x) {
  Node n = key_root(k);
  const int d = (int)k->domain_bits;
  for (int level = 0; level < d; level++)
    n = ggm_step(k, level, n, (uint32_t)(x >> (d - 1 - level)) & 1u);
  return leaf_mask(k, n);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Streaming 16-byte load that is not kept in the caches (global_load_dwordx4 nt).
__device__ __forceinline__ uint4 load_nt(const uint4* p) {
  const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void xor_masked(uint4& acc, const uint4& v,
                                           const uint4& m) {
  acc.x ^= v.x & m.x;
  acc.y ^= v.y & m.y;
  acc.z ^= v.z & m.z;
  acc.w ^= v.w & m.w;
}

__device__ __forceinline__ uint4 xor4(const uint4& a, const uint4& b) {
  return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w);
}

}  // namespace mg
