// Microbenchmarks that set the VALU roofline of the scan kernels on gfx950.
//
// prg_bench   ChaCha8 GGM node expansions with nothing else: each lane walks
//             a chain of `iters` expansions in registers (the node routine of
//             the scan kernels, correction words included), so the rate is
//             the PRG ceiling of every kernel built on it.
// valu_bench  Plain int32 VALU throughput: eight independent chains of
//             v_add_u32 / v_xor_b32 per lane, written as inline asm so each
//             counted op is exactly one instruction. ops / s over the whole
//             device, divided by 256 CUs x 64 lanes, is the issue clock.
//
// C ABI (libmg_microbench.so, measurement only; no counterpart in the
// reference, whose PRG bench is the CPU criterion bench
// crates/morphogen-gpu-dpf/benches/chacha_prg_bench.rs):
//   int mgb_prg(int device, uint32_t blocks, uint32_t iters, double* ms,
//               uint32_t* checksum)           (the kernels' expansion)
//   int mgb_prg2(int device, uint32_t ordered, uint32_t blocks, uint32_t iters,
//                double* ms, uint32_t* checksum)  (routine 0-5 of mg_prg_variants.hpp)
//   int mgb_valu(int device, uint32_t blocks, uint32_t iters, double* ms,
//                uint32_t* checksum)
// Both launch `blocks` blocks of 256 threads, time the launch with HIP events
// on their own stream, and return 0 or the HIP error code.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "mg_prg_variants.hpp"

namespace mg {

template <int kVariant>
__global__ __launch_bounds__(256) void prg_bench(const DevKey* __restrict__ key, uint32_t iters,
                                                 uint32_t* __restrict__ out) {
  const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
  Node n;
  n.s0 = gid * 0x9E3779B9u;
  n.s1 = gid ^ 0x85EBCA6Bu;
  n.s2 = gid + 0xC2B2AE35u;
  n.s3 = ~gid;
  n.t = gid & 1u;
  uint32_t acc = 0;
  for (uint32_t i = 0; i < iters; ++i) {
    Node L, R;
    bench_expand_node<kVariant>(key, (int)(i & 15u), n, L, R);
    // The right child stays live (the scan kernels push it to their stack).
    acc ^= R.s0 ^ R.s1 ^ R.s2 ^ R.s3 ^ R.t;
    n = L;
  }
  out[gid] = acc ^ n.s0 ^ n.s1 ^ n.s2 ^ n.s3 ^ n.t;
}

__global__ __launch_bounds__(256) void valu_bench(uint32_t iters, uint32_t* __restrict__ out) {
  const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
  uint32_t x0 = gid, x1 = gid + 1, x2 = gid + 2, x3 = gid + 3;
  uint32_t x4 = gid + 4, x5 = gid + 5, x6 = gid + 6, x7 = gid + 7;
  const uint32_t y = gid * 0x9E3779B9u;
  for (uint32_t i = 0; i < iters; ++i) {
    // 16 instructions, 8 independent chains.
#define MGB_OP(ins, x) asm volatile(ins " %0, %0, %1" : "+v"(x) : "v"(y))
    MGB_OP("v_add_u32", x0); MGB_OP("v_add_u32", x1); MGB_OP("v_add_u32", x2);
    MGB_OP("v_add_u32", x3); MGB_OP("v_add_u32", x4); MGB_OP("v_add_u32", x5);
    MGB_OP("v_add_u32", x6); MGB_OP("v_add_u32", x7);
    MGB_OP("v_xor_b32", x0); MGB_OP("v_xor_b32", x1); MGB_OP("v_xor_b32", x2);
    MGB_OP("v_xor_b32", x3); MGB_OP("v_xor_b32", x4); MGB_OP("v_xor_b32", x5);
    MGB_OP("v_xor_b32", x6); MGB_OP("v_xor_b32", x7);
#undef MGB_OP
  }
  out[gid] = x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7;
}

// One instruction form per OP, 8 independent chains x 2 per iteration, so a
// rate per form (cycles per wave instruction) comes out of ops / s and the
// clock. The forms are the candidates for the ChaCha rotates and adds.
template <int OP>
__device__ __forceinline__ void op_once(uint32_t& x, uint32_t y) {
  if constexpr (OP == 0) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x) : "v"(y));
  else if constexpr (OP == 1) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(x) : "v"(y));
  else if constexpr (OP == 2) asm volatile("v_alignbit_b32 %0, %0, %0, 25" : "+v"(x));
  else if constexpr (OP == 3) asm volatile("v_perm_b32 %0, %0, %0, %1" : "+v"(x) : "s"(0x01000302u));
  else if constexpr (OP == 4) asm volatile("v_lshlrev_b32 %0, 7, %0" : "+v"(x));
  else if constexpr (OP == 5) asm volatile("v_bitop3_b32 %0, %0, %1, %1 bitop3:0x78" : "+v"(x) : "v"(y));
  else if constexpr (OP == 6) asm volatile("v_add3_u32 %0, %0, %1, %1" : "+v"(x) : "v"(y));
  else if constexpr (OP == 7) asm volatile("v_xad_u32 %0, %0, %1, %1" : "+v"(x) : "v"(y));
  else if constexpr (OP == 8) asm volatile("v_lshl_or_b32 %0, %0, 7, %1" : "+v"(x) : "v"(y));
  else if constexpr (OP == 9) asm volatile("v_alignbyte_b32 %0, %0, %0, 2" : "+v"(x));
  else if constexpr (OP == 10) asm volatile("v_add_u32_e64 %0, %0, %1" : "+v"(x) : "v"(y));
  else if constexpr (OP == 11) asm volatile("v_lshrrev_b32 %0, 25, %0" : "+v"(x));
  else if constexpr (OP == 12) asm volatile("v_or_b32 %0, %0, %1" : "+v"(x) : "v"(y));
  else if constexpr (OP == 13) asm volatile("v_alignbit_b32 %0, %0, %1, 25" : "+v"(x) : "v"(y));
  // Halves swapped by a packed 16-bit add of zero: a rotate by 16.
  else if constexpr (OP == 14)
    asm volatile("v_pk_add_u16 %0, %0, %1 op_sel:[1,0] op_sel_hi:[0,0]" : "+v"(x) : "v"(0u));
  else if constexpr (OP == 15) asm volatile("v_pk_add_u16 %0, %0, %1" : "+v"(x) : "v"(y));
  else if constexpr (OP == 16)
    asm volatile("v_xor_b32_sdwa %0, %0, %1 dst_sel:WORD_0 dst_unused:UNUSED_PAD "
                 "src0_sel:WORD_1 src1_sel:WORD_1" : "+v"(x) : "v"(y));
  else if constexpr (OP == 17)
    asm volatile("v_xor_b32_sdwa %0, %0, %1 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE "
                 "src0_sel:WORD_0 src1_sel:WORD_0" : "+v"(x) : "v"(y));
}

template <int OP>
__global__ __launch_bounds__(256) void op_bench(uint32_t iters, uint32_t* __restrict__ out) {
  const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
  uint32_t x0 = gid, x1 = gid + 1, x2 = gid + 2, x3 = gid + 3;
  uint32_t x4 = gid + 4, x5 = gid + 5, x6 = gid + 6, x7 = gid + 7;
  const uint32_t y = gid * 0x9E3779B9u;
  for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < 2; r++) {
      op_once<OP>(x0, y); op_once<OP>(x1, y); op_once<OP>(x2, y); op_once<OP>(x3, y);
      op_once<OP>(x4, y); op_once<OP>(x5, y); op_once<OP>(x6, y); op_once<OP>(x7, y);
    }
  }
  out[gid] = x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7;
}

// K interleaved dependent chains, 48 instructions per iteration. MIX 0: all
// v_add_u32; MIX 1: each chain cycles add, xor, alignbit (the ChaCha
// quarter-round's pattern); MIX 2: add, xor only.
template <int K, int MIX, int BODY = 1>
__global__ __launch_bounds__(256) void mix_bench(uint32_t iters, uint32_t* __restrict__ out) {
  const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
  uint32_t x[K];
#pragma unroll
  for (int c = 0; c < K; c++) x[c] = gid + (uint32_t)c;
  const uint32_t y = gid * 0x9E3779B9u;
  for (uint32_t i = 0; i < iters; i += BODY) {
#pragma unroll
    for (int r = 0; r < 48 / K * BODY; r++) {
#pragma unroll
      for (int c = 0; c < K; c++) {
        const int kind = MIX == 0 ? 0 : (MIX == 1 ? r % 3 : r % 2);
        if (kind == 0) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x[c]) : "v"(y));
        else if (kind == 1) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(x[c]) : "v"(y));
        else asm volatile("v_alignbit_b32 %0, %0, %0, 25" : "+v"(x[c]));
      }
    }
  }
  uint32_t a = 0;
#pragma unroll
  for (int c = 0; c < K; c++) a ^= x[c];
  out[gid] = a;
}

// VGPR banks: 8 chains v40..v47 += v[x + 8] (same bank, x mod 4) or
// v[x + 9] (next bank), fixed registers, 16 adds per iteration.
template <int SAME>
__global__ __launch_bounds__(256) void bank_bench(uint32_t iters, uint32_t* __restrict__ out) {
  const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
  asm volatile(
      "v_mov_b32 v40, %0\n v_mov_b32 v41, %0\n v_mov_b32 v42, %0\n v_mov_b32 v43, %0\n"
      "v_mov_b32 v44, %0\n v_mov_b32 v45, %0\n v_mov_b32 v46, %0\n v_mov_b32 v47, %0\n"
      "v_mov_b32 v48, %0\n v_mov_b32 v49, %0\n v_mov_b32 v50, %0\n v_mov_b32 v51, %0\n"
      "v_mov_b32 v52, %0\n v_mov_b32 v53, %0\n v_mov_b32 v54, %0\n v_mov_b32 v55, %0\n"
      "v_mov_b32 v56, %0\n"
      :: "v"(gid)
      : "v40", "v41", "v42", "v43", "v44", "v45", "v46", "v47", "v48", "v49", "v50", "v51",
        "v52", "v53", "v54", "v55", "v56");
  for (uint32_t i = 0; i < iters; ++i) {
    if constexpr (SAME) {
      asm volatile(
          "v_add_u32 v40, v40, v48\n v_add_u32 v41, v41, v49\n v_add_u32 v42, v42, v50\n"
          "v_add_u32 v43, v43, v51\n v_add_u32 v44, v44, v52\n v_add_u32 v45, v45, v53\n"
          "v_add_u32 v46, v46, v54\n v_add_u32 v47, v47, v55\n"
          "v_add_u32 v40, v40, v48\n v_add_u32 v41, v41, v49\n v_add_u32 v42, v42, v50\n"
          "v_add_u32 v43, v43, v51\n v_add_u32 v44, v44, v52\n v_add_u32 v45, v45, v53\n"
          "v_add_u32 v46, v46, v54\n v_add_u32 v47, v47, v55\n"
          ::: "v40", "v41", "v42", "v43", "v44", "v45", "v46", "v47");
    } else {
      asm volatile(
          "v_add_u32 v40, v40, v49\n v_add_u32 v41, v41, v50\n v_add_u32 v42, v42, v51\n"
          "v_add_u32 v43, v43, v52\n v_add_u32 v44, v44, v53\n v_add_u32 v45, v45, v54\n"
          "v_add_u32 v46, v46, v55\n v_add_u32 v47, v47, v56\n"
          "v_add_u32 v40, v40, v49\n v_add_u32 v41, v41, v50\n v_add_u32 v42, v42, v51\n"
          "v_add_u32 v43, v43, v52\n v_add_u32 v44, v44, v53\n v_add_u32 v45, v45, v54\n"
          "v_add_u32 v46, v46, v55\n v_add_u32 v47, v47, v56\n"
          ::: "v40", "v41", "v42", "v43", "v44", "v45", "v46", "v47");
    }
  }
  uint32_t r;
  asm volatile(
      "v_xor_b32 %0, v40, v41\n v_xor_b32 %0, %0, v42\n v_xor_b32 %0, %0, v43\n"
      "v_xor_b32 %0, %0, v44\n v_xor_b32 %0, %0, v45\n v_xor_b32 %0, %0, v46\n"
      "v_xor_b32 %0, %0, v47\n"
      : "=&v"(r)::"v40", "v41", "v42", "v43", "v44", "v45", "v46", "v47");
  out[gid] = r;
}

// Two registers per chain, both operands freshly written: MIX 0 alternates
// x += z, z ^= x; MIX 1 adds z = alignbit(z, z, 25) after each pair (three
// ops, the quarter-round's shape). K = 8 chains, 48 instructions per iteration.
template <int MIX>
__global__ __launch_bounds__(256) void fresh_bench(uint32_t iters, uint32_t* __restrict__ out) {
  const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
  uint32_t x[8], z[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    x[c] = gid + (uint32_t)c;
    z[c] = gid * 0x9E3779B9u + (uint32_t)c;
  }
  constexpr int kOps = MIX == 0 ? 2 : 3;
  for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < 48 / 8; r++) {
      const int kind = r % kOps;
#pragma unroll
      for (int c = 0; c < 8; c++) {
        if (kind == 0) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x[c]) : "v"(z[c]));
        else if (kind == 1) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(z[c]) : "v"(x[c]));
        else asm volatile("v_alignbit_b32 %0, %0, %0, 25" : "+v"(z[c]));
      }
    }
  }
  uint32_t a = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) a ^= x[c] ^ z[c];
  out[gid] = a;
}

// Eight independent ChaCha quarter-rounds on registers a, b, c, d, each op
// written as its own asm statement. ORDER 0 issues stage k of all eight QRs
// before stage k + 1 (8 independent ops in a row); ORDER 1 issues each QR's
// 12 dependent ops back to back. 96 instructions per iteration.
#define MGB_ADD(x, y) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x) : "v"(y))
#define MGB_XOR(x, y) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(x) : "v"(y))
#define MGB_ROT(x, n) asm volatile("v_alignbit_b32 %0, %0, %0, " #n : "+v"(x))
// Rotate by 16 as a packed 16-bit add of zero with the halves swapped.
#define MGB_ROT16_PK(x) \
  asm volatile("v_pk_add_u16 %0, %0, %1 op_sel:[1,0] op_sel_hi:[0,0]" : "+v"(x) : "v"(0u))
// x = rotr16(x ^ y) as two word-selecting xors: low half from the high halves,
// then the high half from the low halves.
#define MGB_XOR_ROT16_SDWA(x, y)                                                        \
  do {                                                                                  \
    uint32_t t_;                                                                        \
    asm volatile("v_xor_b32_sdwa %0, %1, %2 dst_sel:WORD_0 dst_unused:UNUSED_PAD "      \
                 "src0_sel:WORD_1 src1_sel:WORD_1" : "=v"(t_) : "v"(x), "v"(y));        \
    asm volatile("v_xor_b32_sdwa %0, %1, %2 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE " \
                 "src0_sel:WORD_0 src1_sel:WORD_0" : "+v"(t_) : "v"(x), "v"(y));        \
    x = t_;                                                                             \
  } while (0)
// ORDER 0: op-major; 1: one quarter-round at a time; 2: op-major with the
// rotate by 16 as MGB_ROT16_PK; 3: op-major with xor + rotate by 16 as
// MGB_XOR_ROT16_SDWA. All four compute the same words.
template <int ORDER, int NQ = 8>
__global__ __launch_bounds__(256) void qr_bench(uint32_t iters, uint32_t* __restrict__ out) {
  const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
  uint32_t a[NQ], b[NQ], c[NQ], d[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    a[q] = gid + (uint32_t)q;
    b[q] = gid * 3u + (uint32_t)q;
    c[q] = gid * 5u + (uint32_t)q;
    d[q] = gid * 7u + (uint32_t)q;
  }
  for (uint32_t i = 0; i < iters; ++i) {
    if constexpr (ORDER != 1) {
#pragma unroll
      for (int q = 0; q < NQ; q++) MGB_ADD(a[q], b[q]);
      if constexpr (ORDER == 3) {
#pragma unroll
        for (int q = 0; q < NQ; q++) MGB_XOR_ROT16_SDWA(d[q], a[q]);
      } else {
#pragma unroll
        for (int q = 0; q < NQ; q++) MGB_XOR(d[q], a[q]);
        if constexpr (ORDER == 2) {
#pragma unroll
          for (int q = 0; q < NQ; q++) MGB_ROT16_PK(d[q]);
        } else {
#pragma unroll
          for (int q = 0; q < NQ; q++) MGB_ROT(d[q], 16);
        }
      }
#pragma unroll
      for (int q = 0; q < NQ; q++) MGB_ADD(c[q], d[q]);
#pragma unroll
      for (int q = 0; q < NQ; q++) MGB_XOR(b[q], c[q]);
#pragma unroll
      for (int q = 0; q < NQ; q++) MGB_ROT(b[q], 20);
#pragma unroll
      for (int q = 0; q < NQ; q++) MGB_ADD(a[q], b[q]);
#pragma unroll
      for (int q = 0; q < NQ; q++) MGB_XOR(d[q], a[q]);
#pragma unroll
      for (int q = 0; q < NQ; q++) MGB_ROT(d[q], 24);
#pragma unroll
      for (int q = 0; q < NQ; q++) MGB_ADD(c[q], d[q]);
#pragma unroll
      for (int q = 0; q < NQ; q++) MGB_XOR(b[q], c[q]);
#pragma unroll
      for (int q = 0; q < NQ; q++) MGB_ROT(b[q], 25);
    } else {
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        MGB_ADD(a[q], b[q]); MGB_XOR(d[q], a[q]); MGB_ROT(d[q], 16);
        MGB_ADD(c[q], d[q]); MGB_XOR(b[q], c[q]); MGB_ROT(b[q], 20);
        MGB_ADD(a[q], b[q]); MGB_XOR(d[q], a[q]); MGB_ROT(d[q], 24);
        MGB_ADD(c[q], d[q]); MGB_XOR(b[q], c[q]); MGB_ROT(b[q], 25);
      }
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int q = 0; q < NQ; q++) r ^= a[q] ^ b[q] ^ c[q] ^ d[q];
  out[gid] = r;
}
// Eight quarter-rounds as two groups of four, group B one stage behind
// group A, their ops interleaved one by one, so a rotate of one group sits
// beside an add or xor of the other. 96 instructions per iteration.
template <int K>
__device__ __forceinline__ void qr_stage(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
  if constexpr (K == 0) MGB_ADD(a, b);
  else if constexpr (K == 1) MGB_XOR(d, a);
  else if constexpr (K == 2) MGB_ROT(d, 16);
  else if constexpr (K == 3) MGB_ADD(c, d);
  else if constexpr (K == 4) MGB_XOR(b, c);
  else if constexpr (K == 5) MGB_ROT(b, 20);
  else if constexpr (K == 6) MGB_ADD(a, b);
  else if constexpr (K == 7) MGB_XOR(d, a);
  else if constexpr (K == 8) MGB_ROT(d, 24);
  else if constexpr (K == 9) MGB_ADD(c, d);
  else if constexpr (K == 10) MGB_XOR(b, c);
  else if constexpr (K == 11) MGB_ROT(b, 25);
}

template <int KA>
__device__ __forceinline__ void skew_step(uint32_t (&a)[8], uint32_t (&b)[8], uint32_t (&c)[8],
                                          uint32_t (&d)[8]) {
  // Group A (QRs 0-3) at stage KA, group B (QRs 4-7) at stage KA - 1.
#pragma unroll
  for (int q = 0; q < 4; q++) {
    qr_stage<KA>(a[q], b[q], c[q], d[q]);
    if constexpr (KA >= 1) qr_stage<KA - 1>(a[q + 4], b[q + 4], c[q + 4], d[q + 4]);
  }
}

__global__ __launch_bounds__(256) void qr_skew_bench(uint32_t iters, uint32_t* __restrict__ out) {
  const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
  uint32_t a[8], b[8], c[8], d[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    a[q] = gid + (uint32_t)q;
    b[q] = gid * 3u + (uint32_t)q;
    c[q] = gid * 5u + (uint32_t)q;
    d[q] = gid * 7u + (uint32_t)q;
  }
  for (uint32_t i = 0; i < iters; ++i) {
    skew_step<0>(a, b, c, d);
    skew_step<1>(a, b, c, d);
    skew_step<2>(a, b, c, d);
    skew_step<3>(a, b, c, d);
    skew_step<4>(a, b, c, d);
    skew_step<5>(a, b, c, d);
    skew_step<6>(a, b, c, d);
    skew_step<7>(a, b, c, d);
    skew_step<8>(a, b, c, d);
    skew_step<9>(a, b, c, d);
    skew_step<10>(a, b, c, d);
    skew_step<11>(a, b, c, d);
#pragma unroll
    for (int q = 0; q < 4; q++) qr_stage<11>(a[q + 4], b[q + 4], c[q + 4], d[q + 4]);
  }
  uint32_t r = 0;
#pragma unroll
  for (int q = 0; q < 8; q++) r ^= a[q] ^ b[q] ^ c[q] ^ d[q];
  out[gid] = r;
}

#undef MGB_ADD
#undef MGB_XOR
#undef MGB_ROT

constexpr int kOpForms = 18;
// mgb_mix forms: (K, MIX).
constexpr int kMixForms = 22;

}  // namespace mg

namespace {

// Times one launch of `launch` on a fresh stream and XOR-folds the output.
template <class F>
int timed(int device, uint32_t blocks, F launch, double* ms, uint32_t* checksum) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return (int)e;
  const size_t n = (size_t)blocks * 256;
  uint32_t* out = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  std::vector<uint32_t> host(n);
  float t = 0;
  if ((e = hipMalloc((void**)&out, n * 4)) != hipSuccess) return (int)e;
  if ((e = hipStreamCreate(&s)) == hipSuccess && (e = hipEventCreate(&a)) == hipSuccess &&
      (e = hipEventCreate(&b)) == hipSuccess && (e = hipEventRecord(a, s)) == hipSuccess &&
      (e = launch(out, s)) == hipSuccess && (e = hipEventRecord(b, s)) == hipSuccess &&
      (e = hipEventSynchronize(b)) == hipSuccess && (e = hipEventElapsedTime(&t, a, b)) == hipSuccess &&
      (e = hipMemcpy(host.data(), out, n * 4, hipMemcpyDeviceToHost)) == hipSuccess) {
    uint32_t x = 0;
    for (uint32_t v : host) x ^= v;
    *checksum = x;
    *ms = t;
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(out);
  return (int)e;
}

}  // namespace

extern "C" {

// ordered: the node routine, 0 to mg::kBenchVariants - 1 (mg_prg_variants.hpp).
int mgb_prg2(int device, uint32_t ordered, uint32_t blocks, uint32_t iters, double* ms,
             uint32_t* checksum) {
  if (!ms || !checksum || !blocks || ordered >= (uint32_t)mg::kBenchVariants)
    return (int)hipErrorInvalidValue;
  // A key with every correction word set, so half the nodes (t == 1) apply it.
  mg::DevKey hk{};
  for (int l = 0; l < mg::kMaxDomainBits; l++) {
    for (int w = 0; w < 4; w++) hk.cw_seed[l][w] = 0x01234567u * (uint32_t)(l * 4 + w + 1);
    hk.cw_t[l] = 0x0101u;
  }
  hk.domain_bits = 16;
  mg::DevKey* dk = nullptr;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipMalloc((void**)&dk, sizeof hk);
  if (e == hipSuccess) e = hipMemcpy(dk, &hk, sizeof hk, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (dk) (void)hipFree(dk);
    return (int)e;
  }
  const int rc = timed(
      device, blocks,
      [&](uint32_t* out, hipStream_t s) {
        switch (ordered) {
#define MGB_PRG(V) \
  case V: hipLaunchKernelGGL(mg::prg_bench<V>, dim3(blocks), dim3(256), 0, s, dk, iters, out); break;
          MGB_PRG(0) MGB_PRG(1) MGB_PRG(2) MGB_PRG(3) MGB_PRG(4) MGB_PRG(5)
#undef MGB_PRG
        }
        static_assert(mg::kBenchVariants == 6, "one case per routine");
        return hipGetLastError();
      },
      ms, checksum);
  (void)hipFree(dk);
  return rc;
}

int mgb_prg(int device, uint32_t blocks, uint32_t iters, double* ms, uint32_t* checksum) {
  return mgb_prg2(device, (uint32_t)mg::kBenchKernelVariant, blocks, iters, ms, checksum);
}

int mgb_valu(int device, uint32_t blocks, uint32_t iters, double* ms, uint32_t* checksum) {
  if (!ms || !checksum || !blocks) return (int)hipErrorInvalidValue;
  return timed(
      device, blocks,
      [&](uint32_t* out, hipStream_t s) {
        hipLaunchKernelGGL(mg::valu_bench, dim3(blocks), dim3(256), 0, s, iters, out);
        return hipGetLastError();
      },
      ms, checksum);
}

// Form `op` (0 .. 17, see op_once) x 16 per lane per iteration.
int mgb_op(int device, uint32_t op, uint32_t blocks, uint32_t iters, double* ms,
           uint32_t* checksum) {
  if (!ms || !checksum || !blocks || op >= (uint32_t)mg::kOpForms)
    return (int)hipErrorInvalidValue;
  return timed(
      device, blocks,
      [&](uint32_t* out, hipStream_t s) {
        switch (op) {
#define MGB_CASE(K) \
  case K: hipLaunchKernelGGL(mg::op_bench<K>, dim3(blocks), dim3(256), 0, s, iters, out); break;
          MGB_CASE(0) MGB_CASE(1) MGB_CASE(2) MGB_CASE(3) MGB_CASE(4) MGB_CASE(5) MGB_CASE(6)
          MGB_CASE(7) MGB_CASE(8) MGB_CASE(9) MGB_CASE(10) MGB_CASE(11) MGB_CASE(12)
          MGB_CASE(13) MGB_CASE(14) MGB_CASE(15) MGB_CASE(16) MGB_CASE(17)
#undef MGB_CASE
        }
        return hipGetLastError();
      },
      ms, checksum);
}

// Dependency pattern `form` (0 .. 9): K = 1, 2, 4, 8, 16 chains of adds (0-4),
// K = 4, 8, 16 chains of add/xor/alignbit (5-7), K = 4, 8 chains of add/xor
// (8-9); 48 instructions per lane per iteration. Forms 10 and 11: 8 chains of
// adds whose second operand sits in the same VGPR bank (10) or the next (11),
// also 48 adds per iteration. Forms 12 and 13: forms 6 and 9 with the loop
// body unrolled 15 times (720 instructions, the size of the PRG loop); iters
// must be a multiple of 15. Forms 14 and 15: fresh_bench<0>, <1>. Forms 16 and
// 17: qr_bench<0>, <1> (iters / 2 iterations of 96). Form 18: 16 quarter-rounds
// op-major (iters / 4 iterations of 192). Form 19: 8 quarter-rounds in two
// groups skewed by one stage (iters / 2 iterations of 96). Forms 20 and 21:
// qr_bench<2>, <3>, form 16 with the rotate by 16 as a packed 16-bit add (20)
// or xor + rotate by 16 as two word-selecting SDWA xors (21).
int mgb_mix(int device, uint32_t form, uint32_t blocks, uint32_t iters, double* ms,
            uint32_t* checksum) {
  if (!ms || !checksum || !blocks || form >= (uint32_t)mg::kMixForms)
    return (int)hipErrorInvalidValue;
  return timed(
      device, blocks,
      [&](uint32_t* out, hipStream_t s) {
        const dim3 g(blocks), b(256);
        switch (form) {
          case 0: hipLaunchKernelGGL((mg::mix_bench<1, 0>), g, b, 0, s, iters, out); break;
          case 1: hipLaunchKernelGGL((mg::mix_bench<2, 0>), g, b, 0, s, iters, out); break;
          case 2: hipLaunchKernelGGL((mg::mix_bench<4, 0>), g, b, 0, s, iters, out); break;
          case 3: hipLaunchKernelGGL((mg::mix_bench<8, 0>), g, b, 0, s, iters, out); break;
          case 4: hipLaunchKernelGGL((mg::mix_bench<16, 0>), g, b, 0, s, iters, out); break;
          case 5: hipLaunchKernelGGL((mg::mix_bench<4, 1>), g, b, 0, s, iters, out); break;
          case 6: hipLaunchKernelGGL((mg::mix_bench<8, 1>), g, b, 0, s, iters, out); break;
          case 7: hipLaunchKernelGGL((mg::mix_bench<16, 1>), g, b, 0, s, iters, out); break;
          case 8: hipLaunchKernelGGL((mg::mix_bench<4, 2>), g, b, 0, s, iters, out); break;
          case 9: hipLaunchKernelGGL((mg::mix_bench<8, 2>), g, b, 0, s, iters, out); break;
          case 10: hipLaunchKernelGGL((mg::bank_bench<1>), g, b, 0, s, iters * 3, out); break;
          case 11: hipLaunchKernelGGL((mg::bank_bench<0>), g, b, 0, s, iters * 3, out); break;
          case 12: hipLaunchKernelGGL((mg::mix_bench<8, 1, 15>), g, b, 0, s, iters, out); break;
          case 13: hipLaunchKernelGGL((mg::mix_bench<8, 2, 15>), g, b, 0, s, iters, out); break;
          case 14: hipLaunchKernelGGL((mg::fresh_bench<0>), g, b, 0, s, iters, out); break;
          case 15: hipLaunchKernelGGL((mg::fresh_bench<1>), g, b, 0, s, iters, out); break;
          // 96 instructions per qr_bench iteration: half the iterations.
          case 16: hipLaunchKernelGGL((mg::qr_bench<0>), g, b, 0, s, iters / 2, out); break;
          case 17: hipLaunchKernelGGL((mg::qr_bench<1>), g, b, 0, s, iters / 2, out); break;
          // 16 quarter-rounds op-major: 192 instructions per iteration.
          case 18: hipLaunchKernelGGL((mg::qr_bench<0, 16>), g, b, 0, s, iters / 4, out); break;
          case 19: hipLaunchKernelGGL(mg::qr_skew_bench, g, b, 0, s, iters / 2, out); break;
          case 20: hipLaunchKernelGGL((mg::qr_bench<2>), g, b, 0, s, iters / 2, out); break;
          case 21: hipLaunchKernelGGL((mg::qr_bench<3>), g, b, 0, s, iters / 2, out); break;
        }
        return hipGetLastError();
      },
      ms, checksum);
}

}  // extern "C"
