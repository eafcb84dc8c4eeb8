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
// low bits of words 0 and 1, so its ops that feed nothing else are dropped.
// Rounds 2-8 are issued in a fixed order: each of the 12
// stages of a quarter-round is issued for all eight quarter-rounds of the
// round (4 of each block) before the next stage, so consecutive instructions
// never depend on each other. gfx950 issues v_add_u32 / v_xor_b32 faster than
// one wave instruction per 4 cycles only when the next instruction does not
// wait for them; the compiler's own order keeps a quarter-round's dependent
// ops close together. Each op is its own asm statement (volatile asm keeps
// the order) in SSA form, so words the two blocks still share need no copy.
// Round 1 stays in C: there the two blocks share three column quarter-rounds
// and the compiler folds the zero words. In round 8 block 1 keeps only what
// its words 0 and 1 need.
__device__ __forceinline__ uint32_t op_add(uint32_t a, uint32_t b) {
  uint32_t r;
  asm volatile("v_add_u32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ uint32_t op_xor(uint32_t a, uint32_t b) {
  uint32_t r;
  asm volatile("v_xor_b32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
// rotl(x, 32 - N) = alignbit(x, x, N)
#define MG_ROTR_ASM(N)                                                         \
  __device__ __forceinline__ uint32_t op_rotr##N(uint32_t a) {                 \
    uint32_t r;                                                                \
    asm volatile("v_alignbit_b32 %0, %1, %1, " #N : "=v"(r) : "v"(a));         \
    return r;                                                                  \
  }
MG_ROTR_ASM(16)
MG_ROTR_ASM(20)
MG_ROTR_ASM(24)
MG_ROTR_ASM(25)
#undef MG_ROTR_ASM

// An s_nop of N wait states (none for N = 0).
template <int N>
__device__ __forceinline__ void stage_nop() {
  static_assert(N >= 0 && N <= 16, "s_nop of 0-16 wait states");
  if constexpr (N > 0) asm volatile("s_nop %0" ::"i"(N - 1));
}

// Stages 1..LAST of NQ quarter-rounds (a[q], b[q], c[q], d[q]) in op-major
// order. NA / NX / NR > 0: an s_nop of that many wait states before each add /
// xor / rotate stage, on top of the one the compiler puts before a stage whose
// ops read the previous stage's results.
template <int NQ, int LAST, int NA = 0, int NR = 0, int NX = NA>
__device__ __forceinline__ void qr_stages(uint32_t* a[], uint32_t* b[], uint32_t* c[],
                                          uint32_t* d[]) {
#define MG_STAGE(K, EXPR)                                                      \
  if (LAST >= K) {                                                             \
    stage_nop<(K % 3 == 0) ? NR : (K % 3 == 1) ? NA : NX>();                   \
    _Pragma("unroll") for (int q = 0; q < NQ; q++) { EXPR; }                   \
  }
  MG_STAGE(1, *a[q] = op_add(*a[q], *b[q]))
  MG_STAGE(2, *d[q] = op_xor(*d[q], *a[q]))
  MG_STAGE(3, *d[q] = op_rotr16(*d[q]))
  MG_STAGE(4, *c[q] = op_add(*c[q], *d[q]))
  MG_STAGE(5, *b[q] = op_xor(*b[q], *c[q]))
  MG_STAGE(6, *b[q] = op_rotr20(*b[q]))
  MG_STAGE(7, *a[q] = op_add(*a[q], *b[q]))
  MG_STAGE(8, *d[q] = op_xor(*d[q], *a[q]))
  MG_STAGE(9, *d[q] = op_rotr24(*d[q]))
  MG_STAGE(10, *c[q] = op_add(*c[q], *d[q]))
  MG_STAGE(11, *b[q] = op_xor(*b[q], *c[q]))
  MG_STAGE(12, *b[q] = op_rotr25(*b[q]))
#undef MG_STAGE
}

__device__ __forceinline__ void prg_expand_ordered(uint32_t s0, uint32_t s1, uint32_t s2,
                                                   uint32_t s3, uint32_t l[4], uint32_t r[4],
                                                   uint32_t& lt, uint32_t& rt) {
  uint32_t x0 = kC0, x1 = kC1, x2 = kC2, x3 = kC3;
  uint32_t x4 = s0, x5 = s1, x6 = s2, x7 = s3;
  uint32_t x8 = s0, x9 = s1, x10 = s2, x11 = s3;
  uint32_t x12 = 0, x13 = 0, x14 = 0, x15 = 0;
  uint32_t y0 = kC0, y4 = s0, y8 = s0, y12 = 1;
  // Round 1 (columns): block 1 differs from block 0 only in column 0.
  MG_QR(x0, x4, x8, x12);
  MG_QR(x1, x5, x9, x13);
  MG_QR(x2, x6, x10, x14);
  MG_QR(x3, x7, x11, x15);
  MG_QR(y0, y4, y8, y12);
  uint32_t y1 = x1, y5 = x5, y9 = x9, y13 = x13;
  uint32_t y2 = x2, y6 = x6, y10 = x10, y14 = x14;
  uint32_t y3 = x3, y7 = x7, y11 = x11, y15 = x15;
  // Rounds 2..7: diagonal, column, diagonal, column, diagonal, column.
#pragma unroll
  for (int rd = 2; rd <= 7; rd++) {
    if (rd % 2 == 0) {
      uint32_t* a[8] = {&x0, &x1, &x2, &x3, &y0, &y1, &y2, &y3};
      uint32_t* b[8] = {&x5, &x6, &x7, &x4, &y5, &y6, &y7, &y4};
      uint32_t* c[8] = {&x10, &x11, &x8, &x9, &y10, &y11, &y8, &y9};
      uint32_t* d[8] = {&x15, &x12, &x13, &x14, &y15, &y12, &y13, &y14};
      qr_stages<8, 12>(a, b, c, d);
    } else {
      uint32_t* a[8] = {&x0, &x1, &x2, &x3, &y0, &y1, &y2, &y3};
      uint32_t* b[8] = {&x4, &x5, &x6, &x7, &y4, &y5, &y6, &y7};
      uint32_t* c[8] = {&x8, &x9, &x10, &x11, &y8, &y9, &y10, &y11};
      uint32_t* d[8] = {&x12, &x13, &x14, &x15, &y12, &y13, &y14, &y15};
      if (rd < 7) {
        qr_stages<8, 12>(a, b, c, d);
      } else {
        // Round 7 without block 1's dead tail (as MG_A_ROUND7N): its column 0
        // stops after stage 9, its column 3 after stage 10.
        qr_stages<8, 9>(a, b, c, d);
        uint32_t* c7[7] = {&x8, &x9, &x10, &x11, &y9, &y10, &y11};
        uint32_t* d7[7] = {&x12, &x13, &x14, &x15, &y13, &y14, &y15};
#pragma unroll
        for (int q = 0; q < 7; q++) *c7[q] = op_add(*c7[q], *d7[q]);
        uint32_t* b6[6] = {&x4, &x5, &x6, &x7, &y5, &y6};
        uint32_t* c6[6] = {&x8, &x9, &x10, &x11, &y9, &y10};
#pragma unroll
        for (int q = 0; q < 6; q++) *b6[q] = op_xor(*b6[q], *c6[q]);
#pragma unroll
        for (int q = 0; q < 6; q++) *b6[q] = op_rotr25(*b6[q]);
      }
    }
  }
  // Round 8 (diagonal): block 0 in full; block 1's quarter-rounds through
  // words 0 and 1 up to stage 7, which sets them.
  {
    uint32_t* a[6] = {&x0, &x1, &x2, &x3, &y0, &y1};
    uint32_t* b[6] = {&x5, &x6, &x7, &x4, &y5, &y6};
    uint32_t* c[6] = {&x10, &x11, &x8, &x9, &y10, &y11};
    uint32_t* d[6] = {&x15, &x12, &x13, &x14, &y15, &y12};
    qr_stages<6, 7>(a, b, c, d);
    uint32_t* a4[4] = {&x0, &x1, &x2, &x3};
    uint32_t* b4[4] = {&x5, &x6, &x7, &x4};
    uint32_t* c4[4] = {&x10, &x11, &x8, &x9};
    uint32_t* d4[4] = {&x15, &x12, &x13, &x14};
    // Stages 8..12 of block 0: run stages 1..12 would redo 1..7, so the tail
    // is written out.
#pragma unroll
    for (int q = 0; q < 4; q++) *d4[q] = op_xor(*d4[q], *a4[q]);
#pragma unroll
    for (int q = 0; q < 4; q++) *d4[q] = op_rotr24(*d4[q]);
#pragma unroll
    for (int q = 0; q < 4; q++) *c4[q] = op_add(*c4[q], *d4[q]);
#pragma unroll
    for (int q = 0; q < 4; q++) *b4[q] = op_xor(*b4[q], *c4[q]);
#pragma unroll
    for (int q = 0; q < 4; q++) *b4[q] = op_rotr25(*b4[q]);
  }
  l[0] = x0 + kC0; l[1] = x1 + kC1; l[2] = x2 + kC2; l[3] = x3 + kC3;
  r[0] = x4 + s0; r[1] = x5 + s1; r[2] = x6 + s2; r[3] = x7 + s3;
  lt = (y0 + kC0) & 1u;
  rt = (y1 + kC1) & 1u;
}

// Rounds as single asm statements. The compiler's hazard check cannot see
// into inline asm and puts an s_nop between any two asm statements where the
// second reads a register the first wrote (nothing else lies between them to
// count as a wait state): once per op-major stage, 75 per node with one asm
// statement per op. Inside one statement the hardware's own interlocks order
// the ops. Operands %0-%15 are block 0's words, %16-%31 block 1's.
#define MG_A_S1(a, b, c, d) "v_add_u32 %" #a ", %" #a ", %" #b "\n\t"
#define MG_A_S2(a, b, c, d) "v_xor_b32 %" #d ", %" #d ", %" #a "\n\t"
#define MG_A_S3(a, b, c, d) "v_alignbit_b32 %" #d ", %" #d ", %" #d ", 16\n\t"
#define MG_A_S4(a, b, c, d) "v_add_u32 %" #c ", %" #c ", %" #d "\n\t"
#define MG_A_S5(a, b, c, d) "v_xor_b32 %" #b ", %" #b ", %" #c "\n\t"
#define MG_A_S6(a, b, c, d) "v_alignbit_b32 %" #b ", %" #b ", %" #b ", 20\n\t"
#define MG_A_S9(a, b, c, d) "v_alignbit_b32 %" #d ", %" #d ", %" #d ", 24\n\t"
#define MG_A_S12(a, b, c, d) "v_alignbit_b32 %" #b ", %" #b ", %" #b ", 25\n\t"
// The eight quarter-rounds of a column / diagonal round over both blocks.
#define MG_A_COLS(S)                                                               \
  S(0, 4, 8, 12) S(1, 5, 9, 13) S(2, 6, 10, 14) S(3, 7, 11, 15) S(16, 20, 24, 28) \
  S(17, 21, 25, 29) S(18, 22, 26, 30) S(19, 23, 27, 31)
#define MG_A_DIAGS(S)                                                              \
  S(0, 5, 10, 15) S(1, 6, 11, 12) S(2, 7, 8, 13) S(3, 4, 9, 14) S(16, 21, 26, 31) \
  S(17, 22, 27, 28) S(18, 23, 24, 29) S(19, 20, 25, 30)
// A / X / R: text put before each add / xor / rotate stage (empty, or an
// s_nop).
#define MG_A_ROUNDN(QRS, A, X, R)                                                       \
  A QRS(MG_A_S1) X QRS(MG_A_S2) R QRS(MG_A_S3) A QRS(MG_A_S4) X QRS(MG_A_S5) R          \
  QRS(MG_A_S6) A QRS(MG_A_S1) X QRS(MG_A_S2) R QRS(MG_A_S9) A QRS(MG_A_S4) X            \
  QRS(MG_A_S5) R QRS(MG_A_S12)
// Round 7 (columns) without block 1's dead tail: after round 8 block 1 only
// needs its words 0 and 1, whose round-8 quarter-rounds read words 0, 5, 10,
// 15, 1, 6, 11 and 12 of round 7. Its column 0 (0, 4, 8, 12) then only needs
// words 0 and 12 (stages 1-9), column 3 (3, 7, 11, 15) words 11 and 15
// (stages 1-10): 5 ops fewer.
#define MG_A_COLS_7(S)                                                             \
  S(0, 4, 8, 12) S(1, 5, 9, 13) S(2, 6, 10, 14) S(3, 7, 11, 15) S(17, 21, 25, 29) \
  S(18, 22, 26, 30) S(19, 23, 27, 31)
#define MG_A_COLS_6(S) \
  S(0, 4, 8, 12) S(1, 5, 9, 13) S(2, 6, 10, 14) S(3, 7, 11, 15) S(17, 21, 25, 29) S(18, 22, 26, 30)
#define MG_A_ROUND7N(A, X, R)                                                              \
  A MG_A_COLS(MG_A_S1) X MG_A_COLS(MG_A_S2) R MG_A_COLS(MG_A_S3) A MG_A_COLS(MG_A_S4) X    \
  MG_A_COLS(MG_A_S5) R MG_A_COLS(MG_A_S6) A MG_A_COLS(MG_A_S1) X MG_A_COLS(MG_A_S2) R      \
  MG_A_COLS(MG_A_S9) A MG_A_COLS_7(MG_A_S4) X MG_A_COLS_6(MG_A_S5) R MG_A_COLS_6(MG_A_S12)

// Round 8: block 0's diagonals in full; block 1's quarter-rounds through its
// words 0 and 1, (0, 5, 10, 15) and (1, 6, 11, 12), as %16-%19 and %20-%23,
// up to stage 7, which sets them.
#define MG_A_DIAGS6(S) \
  S(0, 5, 10, 15) S(1, 6, 11, 12) S(2, 7, 8, 13) S(3, 4, 9, 14) S(16, 17, 18, 19) S(20, 21, 22, 23)
#define MG_A_DIAGS4(S) S(0, 5, 10, 15) S(1, 6, 11, 12) S(2, 7, 8, 13) S(3, 4, 9, 14)
#define MG_A_ROUND8N(A, X, R)                                                               \
  A MG_A_DIAGS6(MG_A_S1) X MG_A_DIAGS6(MG_A_S2) R MG_A_DIAGS6(MG_A_S3) A                    \
  MG_A_DIAGS6(MG_A_S4) X MG_A_DIAGS6(MG_A_S5) R MG_A_DIAGS6(MG_A_S6) A MG_A_DIAGS6(MG_A_S1) \
  X MG_A_DIAGS4(MG_A_S2) R MG_A_DIAGS4(MG_A_S9) A MG_A_DIAGS4(MG_A_S4) X                    \
  MG_A_DIAGS4(MG_A_S5) R MG_A_DIAGS4(MG_A_S12)
#define MG_A_X32(x, y)                                                                 \
  "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]), "+v"(x[4]), "+v"(x[5]), "+v"(x[6]),  \
  "+v"(x[7]), "+v"(x[8]), "+v"(x[9]), "+v"(x[10]), "+v"(x[11]), "+v"(x[12]),            \
  "+v"(x[13]), "+v"(x[14]), "+v"(x[15]), "+v"(y[0]), "+v"(y[1]), "+v"(y[2]), "+v"(y[3]), \
  "+v"(y[4]), "+v"(y[5]), "+v"(y[6]), "+v"(y[7]), "+v"(y[8]), "+v"(y[9]), "+v"(y[10]),  \
  "+v"(y[11]), "+v"(y[12]), "+v"(y[13]), "+v"(y[14]), "+v"(y[15])

// As prg_expand_ordered (same op order), with rounds 3-8 as one asm statement
// each: 6 hazard nops for those rounds instead of about 63. Round 2 keeps one
// statement per op: its in-place form would need copies of the twelve words
// the blocks still share after round 1. NA / NX / NR: wait states of the s_nop
// put before each add / xor / rotate stage of rounds 3-8 (0 for all three: no
// s_nop; otherwise each at least 1; 1, 1, 1 is the instruction stream of
// prg_expand_ordered). R2N: space round 2's stages the same way (NA - 1,
// NX - 1, NR - 1 wait states on top of the compiler's one).
template <int NA = 0, int NX = 0, int NR = 0, bool R2N = false>
__device__ __forceinline__ void prg_expand_rounds(uint32_t s0, uint32_t s1, uint32_t s2,
                                                  uint32_t s3, uint32_t l[4], uint32_t r[4],
                                                  uint32_t& lt, uint32_t& rt) {
  static_assert((NA == 0 && NX == 0 && NR == 0) || (NA >= 1 && NX >= 1 && NR >= 1 &&
                                                    NA <= 16 && NX <= 16 && NR <= 16),
                "s_nop spacing of 1-16 wait states, or none");
  uint32_t x0 = kC0, x1 = kC1, x2 = kC2, x3 = kC3;
  uint32_t x4 = s0, x5 = s1, x6 = s2, x7 = s3;
  uint32_t x8 = s0, x9 = s1, x10 = s2, x11 = s3;
  uint32_t x12 = 0, x13 = 0, x14 = 0, x15 = 0;
  uint32_t y0 = kC0, y4 = s0, y8 = s0, y12 = 1;
  // Round 1 (columns): block 1 differs from block 0 only in column 0.
  MG_QR(x0, x4, x8, x12);
  MG_QR(x1, x5, x9, x13);
  MG_QR(x2, x6, x10, x14);
  MG_QR(x3, x7, x11, x15);
  MG_QR(y0, y4, y8, y12);
  uint32_t y1 = x1, y5 = x5, y9 = x9, y13 = x13;
  uint32_t y2 = x2, y6 = x6, y10 = x10, y14 = x14;
  uint32_t y3 = x3, y7 = x7, y11 = x11, y15 = x15;
  {  // Round 2 (diagonals), one statement per op.
    uint32_t* a[8] = {&x0, &x1, &x2, &x3, &y0, &y1, &y2, &y3};
    uint32_t* b[8] = {&x5, &x6, &x7, &x4, &y5, &y6, &y7, &y4};
    uint32_t* c[8] = {&x10, &x11, &x8, &x9, &y10, &y11, &y8, &y9};
    uint32_t* d[8] = {&x15, &x12, &x13, &x14, &y15, &y12, &y13, &y14};
    if constexpr (R2N && NA > 0)
      qr_stages<8, 12, NA - 1, NR - 1, NX - 1>(a, b, c, d);
    else
      qr_stages<8, 12>(a, b, c, d);
  }
  uint32_t x[16] = {x0, x1, x2, x3, x4, x5, x6, x7, x8, x9, x10, x11, x12, x13, x14, x15};
  uint32_t y[16] = {y0, y1, y2, y3, y4, y5, y6, y7, y8, y9, y10, y11, y12, y13, y14, y15};
  uint32_t z[8];
  if constexpr (NA == 0) {
    asm volatile(MG_A_ROUNDN(MG_A_COLS, "", "", "") : MG_A_X32(x, y));   // round 3
    asm volatile(MG_A_ROUNDN(MG_A_DIAGS, "", "", "") : MG_A_X32(x, y));  // round 4
    asm volatile(MG_A_ROUNDN(MG_A_COLS, "", "", "") : MG_A_X32(x, y));   // round 5
    asm volatile(MG_A_ROUNDN(MG_A_DIAGS, "", "", "") : MG_A_X32(x, y));  // round 6
    asm volatile(MG_A_ROUND7N("", "", "") : MG_A_X32(x, y));             // round 7
  } else {
    // The s_nop lengths are operands %32-%34 (immediates).
#define MG_A_R(QRS)                                                                     \
  asm volatile(MG_A_ROUNDN(QRS, "s_nop %32\n\t", "s_nop %33\n\t", "s_nop %34\n\t")      \
               : MG_A_X32(x, y)                                                         \
               : "i"(NA - 1), "i"(NX - 1), "i"(NR - 1))
    MG_A_R(MG_A_COLS);   // round 3
    MG_A_R(MG_A_DIAGS);  // round 4
    MG_A_R(MG_A_COLS);   // round 5
    MG_A_R(MG_A_DIAGS);  // round 6
#undef MG_A_R
    asm volatile(MG_A_ROUND7N("s_nop %32\n\t", "s_nop %33\n\t", "s_nop %34\n\t")  // round 7
                 : MG_A_X32(x, y)
                 : "i"(NA - 1), "i"(NX - 1), "i"(NR - 1));
  }
  z[0] = y[0]; z[1] = y[5]; z[2] = y[10]; z[3] = y[15];
  z[4] = y[1]; z[5] = y[6]; z[6] = y[11]; z[7] = y[12];
#define MG_A_Z24                                                                       \
  "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]), "+v"(x[4]), "+v"(x[5]), "+v"(x[6]),  \
  "+v"(x[7]), "+v"(x[8]), "+v"(x[9]), "+v"(x[10]), "+v"(x[11]), "+v"(x[12]),            \
  "+v"(x[13]), "+v"(x[14]), "+v"(x[15]), "+v"(z[0]), "+v"(z[1]), "+v"(z[2]), "+v"(z[3]), \
  "+v"(z[4]), "+v"(z[5]), "+v"(z[6]), "+v"(z[7])
  if constexpr (NA == 0) {
    asm volatile(MG_A_ROUND8N("", "", "") : MG_A_Z24);
  } else {
    // The s_nop lengths are operands %24-%26.
    asm volatile(MG_A_ROUND8N("s_nop %24\n\t", "s_nop %25\n\t", "s_nop %26\n\t")
                 : MG_A_Z24
                 : "i"(NA - 1), "i"(NX - 1), "i"(NR - 1));
  }
#undef MG_A_Z24
  l[0] = x[0] + kC0; l[1] = x[1] + kC1; l[2] = x[2] + kC2; l[3] = x[3] + kC3;
  r[0] = x[4] + s0; r[1] = x[5] + s1; r[2] = x[6] + s2; r[3] = x[7] + s3;
  lt = (z[0] + kC0) & 1u;
  rt = (z[4] + kC1) & 1u;
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

// The node routines of the kernels. kPrgSpaced, prg_expand_rounds with two
// wait states before each add and xor stage and ten before each rotate stage
// of rounds 2-8, is the lane-DFS and eval kernels' routine: 4.6 % more node
// expansions per second than kPrgOrdered at 4 waves per SIMD, 6.5 % at 5,
// 9.7 % at 8, and in the 2^31 x 32 B, 24-key lane-DFS pass 800.9-802.5 ms
// against 817.3-818.2 ms with six wait states before rotates (twelve:
// 804.9-805.6, fourteen: 831.1-832.0 ms; at eight waves per SIMD ten still wins:
// 790.3-793.7 ms against 792.1-793.9 for twelve and 807-814 for six, eight and
// fourteen). kPrgOrdered is the routine of the stream kernel's waves and of
// split pairs with one DFS wave: one to three DFS waves per SIMD, where the
// spacing s_nops cost more than they save (78,643,200 x 1 KiB, 4 keys:
// 14.40-14.61 against 14.87-15.45 ms; 238,609,294 x 144 B, 1 key: 9.57-9.61
// against 9.83-9.89 ms). Split pairs with two DFS waves use kPrgSpaced.
enum PrgRoutine : int { kPrgOrdered = 1, kPrgSpaced = 2 };
constexpr int kSpaceAdd = 2, kSpaceXor = 2, kSpaceRot = 10;
constexpr int kPrgVariant = kPrgSpaced;
constexpr int kStreamPrgVariant = kPrgOrdered;
constexpr int kSplitPrgVariant = kPrgOrdered;

// Both children of `n` at tree level `level`, correction word applied when
// n.t == 1 (dpf.rs:221-239). `k` is wave-uniform, so the correction word is
// read with scalar loads.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// 16-byte words in LDS / global memory, typed by address space.
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;
typedef __attribute__((address_space(1))) u32x4 glob_u32x4;
typedef __attribute__((address_space(1))) uint32_t glob_u32;

// Key reads go through the constant address space: no kernel writes a key,
// and a wave-uniform read then becomes a scalar load into SGPRs (a divergent
// one stays a vector load). Through plain pointers the compiler loads them
// with vector loads again after every barrier. Kernels whose waves hand
// masks over through LDS keep the vector loads (kScalarKey = false): a wave
// waiting for a scalar load also waits for its own LDS stores, which share
// the same counter.
typedef __attribute__((address_space(4))) const u32x4 key_u32x4;
typedef __attribute__((address_space(4))) const uint32_t key_u32;
__device__ __forceinline__ u32x4 key_load4(const uint32_t* p) {
  return *(key_u32x4*)p;  // an address-space cast, which reinterpret_cast cannot express
}
__device__ __forceinline__ uint32_t key_load(const uint32_t* p) {
  return *(key_u32*)p;
}
template <bool kScalarKey>
__device__ __forceinline__ u32x4 key_read4(const uint32_t* p) {
  if constexpr (kScalarKey) return key_load4(p);
  else return *reinterpret_cast<const u32x4*>(p);
}
template <bool kScalarKey>
__device__ __forceinline__ uint32_t key_read(const uint32_t* p) {
  if constexpr (kScalarKey) return key_load(p);
  else return *p;
}

// The correction step of dpf.rs:221-239 on the PRG outputs l / r (seeds)
// and lt / rt (control bits) of a node whose control bit is t.
template <bool kScalarKey>
__device__ __forceinline__ void correct_children(const DevKey* __restrict__ k, int level,
                                                 uint32_t t, const uint32_t l[4],
                                                 const uint32_t r[4], uint32_t lt, uint32_t rt,
                                                 Node& left, Node& right) {
  const uint32_t m = t_mask(t);
  const u32x4 cw = key_read4<kScalarKey>(k->cw_seed[level]);
  const uint32_t c0 = cw.x, c1 = cw.y, c2 = cw.z, c3 = cw.w;
  const uint32_t ct = key_read<kScalarKey>(&k->cw_t[level]);
  left.s0 = l[0] ^ (c0 & m); left.s1 = l[1] ^ (c1 & m);
  left.s2 = l[2] ^ (c2 & m); left.s3 = l[3] ^ (c3 & m);
  right.s0 = r[0] ^ (c0 & m); right.s1 = r[1] ^ (c1 & m);
  right.s2 = r[2] ^ (c2 & m); right.s3 = r[3] ^ (c3 & m);
  left.t = lt ^ (ct & 0xFFu & m);
  right.t = rt ^ ((ct >> 8) & 0xFFu & m);
}

template <int kRoutine = kPrgVariant, bool kScalarKey = true>
__device__ __forceinline__ void expand_node(const DevKey* __restrict__ k,
                                            int level, const Node& n,
                                            Node& left, Node& right) {
  static_assert(kRoutine == kPrgOrdered || kRoutine == kPrgSpaced, "a kernel node routine");
  uint32_t l[4], r[4], lt, rt;
  if constexpr (kRoutine == kPrgSpaced)
    prg_expand_rounds<kSpaceAdd, kSpaceXor, kSpaceRot, true>(n.s0, n.s1, n.s2, n.s3, l, r,
                                                             lt, rt);
  else
    prg_expand_ordered(n.s0, n.s1, n.s2, n.s3, l, r, lt, rt);
  correct_children<kScalarKey>(k, level, n.t, l, r, lt, rt, left, right);
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
template <bool kScalarKey = true>
__device__ __forceinline__ uint4 leaf_mask(const DevKey* __restrict__ k,
                                           const Node& n) {
  const uint32_t m = t_mask(n.t);
  const u32x4 f = key_read4<kScalarKey>(k->final_cw);
  return make_uint4(n.s0 ^ (f.x & m), n.s1 ^ (f.y & m), n.s2 ^ (f.z & m), n.s3 ^ (f.w & m));
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
