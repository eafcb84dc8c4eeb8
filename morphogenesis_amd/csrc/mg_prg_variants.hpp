// Node routines that only the PRG microbenchmark builds: the reference points
// the kernels' routines are measured against, next to those routines.
//
//   0  compiler-scheduled expansion (prg_expand_compiler_order below)
//   1  op-major, one asm statement per op (kPrgOrdered, the stream and split
//      kernels' DFS waves)
//   2  op-major, rounds 3-8 as one asm statement each, no s_nop spacing
//   3  rounds 2-8 spaced 2 / 2 / 6 wait states before add / xor / rotate stages
//   4  rounds 2-8 spaced 2 / 2 / 10 (kPrgSpaced, the lane-DFS and eval kernels)
//   5  rounds 2-8 spaced 2 / 2 / 14
#pragma once

#include "mg_device.hpp"

namespace mg {

constexpr int kBenchVariants = 6;
constexpr int kBenchKernelVariant = 4;

// PRG expand (chacha_prg.rs:153-178) written as plain C with both blocks in
// full; the compiler drops block 1's unused words, shares its three
// counter-free first-round column quarter-rounds with block 0 and schedules
// the rest itself.
__device__ __forceinline__ void prg_expand_compiler_order(uint32_t s0, uint32_t s1, uint32_t s2,
                                                          uint32_t s3, uint32_t l[4],
                                                          uint32_t r[4], uint32_t& lt,
                                                          uint32_t& rt) {
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

// Both children of `n` with microbenchmark routine V (0 to kBenchVariants - 1).
template <int V>
__device__ __forceinline__ void bench_expand_node(const DevKey* __restrict__ k, int level,
                                                  const Node& n, Node& left, Node& right) {
  static_assert(V >= 0 && V < kBenchVariants, "a microbenchmark node routine");
  if constexpr (V == 1) {
    expand_node<kPrgOrdered>(k, level, n, left, right);
  } else if constexpr (V == kBenchKernelVariant) {
    expand_node<kPrgSpaced>(k, level, n, left, right);
  } else {
    uint32_t l[4], r[4], lt, rt;
    if constexpr (V == 0)
      prg_expand_compiler_order(n.s0, n.s1, n.s2, n.s3, l, r, lt, rt);
    else if constexpr (V == 2)
      prg_expand_rounds<0, 0, 0>(n.s0, n.s1, n.s2, n.s3, l, r, lt, rt);
    else if constexpr (V == 3)
      prg_expand_rounds<2, 2, 6, true>(n.s0, n.s1, n.s2, n.s3, l, r, lt, rt);
    else
      prg_expand_rounds<2, 2, 14, true>(n.s0, n.s1, n.s2, n.s3, l, r, lt, rt);
    correct_children<true>(k, level, n.t, l, r, lt, rt, left, right);
  }
}

}  // namespace mg
