// gfx950 kernels of the fused ChaCha8-DPF masked-XOR scan.
//
// Answer for one key (kernel.rs:1243-1256, 1276-1383 generalised to W-byte
// rows, one DPF leaf per row): ans = XOR_i rows[i] & tile16(mask(i)).
//
// scan_lane_dfs<C>   the hot kernel for rows of C 16-byte chunks (C <= 8).
//                    Each wave owns one key and a contiguous range of tiles;
//                    each lane walks its own 2^kl-leaf subtree depth-first,
//                    so the GGM tree costs one PRG expansion per leaf pair
//                    plus a short per-tile path walk.
// scan_stream<S,G,PK> rows wider than 128 B: same lane-DFS PRG for G keys
//                    per wave; leaf masks go through LDS and each row is read
//                    once, coalesced, for all G keys (PK rows per load when a
//                    row has at most 64 / PK chunks).
// scan_split<S,G,NP,DW> the stream kernel's work split between NP pairs of
//                    waves per block: DW DFS waves make the masks, a streaming
//                    wave folds the rows (528 B - 4 KiB rows, several keys).
// scan_point         point-eval fallback (d = 0, tiny inputs) and the delta
//                    scan: one lane per row, full path walk per row.
// reduce_partials    XOR of per-wave partial answers into the output.
// fill_*             device-side seeded matrix fills.
#include "mg_device.hpp"
#include "mg_kernels.hpp"

namespace mg {

// ------------------------------------------------------------- helpers ----

// Materialises a leaf mask once, so each row chunk folds in with a single
// v_bitop3_b32 instead of re-deriving the mask per chunk.
__device__ __forceinline__ uint4 opaque4(uint4 v) {
  asm("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
  return v;
}

// acc ^ (v & m) as one v_bitop3_b32 (truth table 0x78 over S0 = acc, S1 = v,
// S2 = m). Used where the compiler would otherwise emit v_and + v_xor.
__device__ __forceinline__ void xor_masked3(uint4& acc, const uint4& v, const uint4& m) {
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x78" : "=v"(acc.x) : "v"(acc.x), "v"(v.x), "v"(m.x));
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x78" : "=v"(acc.y) : "v"(acc.y), "v"(v.y), "v"(m.y));
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x78" : "=v"(acc.z) : "v"(acc.z), "v"(v.z), "v"(m.z));
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x78" : "=v"(acc.w) : "v"(acc.w), "v"(v.w), "v"(m.w));
}

__device__ __forceinline__ uint32_t wave_xor(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v ^= __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ uint4 wave_xor4(uint4 v) {
  return make_uint4(wave_xor(v.x), wave_xor(v.y), wave_xor(v.z), wave_xor(v.w));
}

// XOR of v over the lanes l, l ^ from, l ^ 2*from, ... (from a power of two).
__device__ __forceinline__ uint4 xor_lanes_from(uint4 v, int from) {
  for (int off = from; off < 64; off <<= 1) {
    v.x ^= __shfl_xor(v.x, off, 64);
    v.y ^= __shfl_xor(v.y, off, 64);
    v.z ^= __shfl_xor(v.z, off, 64);
    v.w ^= __shfl_xor(v.w, off, 64);
  }
  return v;
}

// Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// The launch has a multiple of 8 blocks; logical block L = (b % 8) * per + b / 8
// gives each XCD a contiguous range of logical blocks, so the waves that share
// a tile range (same part, different keys) run on one XCD and read their rows
// through one L2. Logical waves past `total` are idle.
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t xcd_logical_wave(uint32_t total, uint32_t wpb, uint32_t wib) {
  const uint32_t blocks = (total + wpb - 1) / wpb;
  const uint32_t per = (blocks + kXcds - 1) / kXcds;
  const uint32_t b = blockIdx.x;
  return ((b % kXcds) * per + b / kXcds) * wpb + wib;
}

static uint32_t xcd_grid(uint32_t total, uint32_t wpb) {
  const uint32_t blocks = (total + wpb - 1) / wpb;
  return (blocks + kXcds - 1) / kXcds * kXcds;
}

// A path of wave-uniform nodes, one level per lane: lane L holds level L.
// Reads use readlane with a uniform index; writes are a per-lane select.
struct LanePath {
  uint32_t s0, s1, s2, s3, t;
  __device__ __forceinline__ Node get(int level) const {
    Node n;
    n.s0 = __builtin_amdgcn_readlane(s0, level);
    n.s1 = __builtin_amdgcn_readlane(s1, level);
    n.s2 = __builtin_amdgcn_readlane(s2, level);
    n.s3 = __builtin_amdgcn_readlane(s3, level);
    n.t = __builtin_amdgcn_readlane(t, level);
    return n;
  }
  __device__ __forceinline__ void put(int lane, int level, const Node& n) {
    const bool mine = lane == level;
    s0 = mine ? n.s0 : s0;
    s1 = mine ? n.s1 : s1;
    s2 = mine ? n.s2 : s2;
    s3 = mine ? n.s3 : s3;
    t = mine ? n.t : t;
  }
};

// Walks the wave-uniform tile prefix for `tile` (tb levels), reusing the
// levels shared with the previous tile, then each lane's lb-level path to its
// own subtree root. Returns the lane's subtree root (tree level tb + lb).
__device__ __forceinline__ Node tile_lane_root(const DevKey* __restrict__ k,
                                               LanePath& path, int lane,
                                               uint64_t tile, uint64_t prev,
                                               bool first, int tb, int lb) {
  int start = 0;
  if (!first) {
    const uint64_t diff = tile ^ prev;  // non-zero, below bit tb
    start = tb - (64 - __builtin_clzll(diff));
  }
  Node n = path.get(start);
  for (int lvl = start; lvl < tb; ++lvl) {
    n = ggm_step(k, lvl, n, (uint32_t)(tile >> (tb - 1 - lvl)) & 1u);
    path.put(lane, lvl + 1, n);
  }
  for (int j = 0; j < lb; ++j)
    n = ggm_step(k, tb + j, n, ((uint32_t)lane >> (lb - 1 - j)) & 1u);
  return n;
}

// The same walk with the path kept in global memory instead of registers:
// pseed[l] / pt[l] hold the wave-uniform node at level l (0 .. tb). Every lane
// writes the same words and later reads them back itself, so program order
// alone orders each lane's read after its write. The path is touched only at
// tile starts, so it costs no registers through the ChaCha rounds.
__device__ __forceinline__ Node tile_lane_root_cached(const DevKey* __restrict__ k,
                                                      glob_u32x4* pseed, glob_u32* pt, int lane,
                                                      uint64_t tile, uint64_t prev, bool first,
                                                      int tb, int lb) {
  int start = 0;
  if (!first) {
    const uint64_t diff = tile ^ prev;  // non-zero, below bit tb
    start = tb - (64 - __builtin_clzll(diff));
  }
  Node n;
  if (start == 0) {
    n = key_root(k);
  } else {
    const u32x4 v = pseed[start];
    n.s0 = v.x; n.s1 = v.y; n.s2 = v.z; n.s3 = v.w;
    n.t = pt[start];
  }
  for (int lvl = start; lvl < tb; ++lvl) {
    n = ggm_step(k, lvl, n, (uint32_t)(tile >> (tb - 1 - lvl)) & 1u);
    pseed[lvl + 1] = u32x4{n.s0, n.s1, n.s2, n.s3};
    pt[lvl + 1] = n.t;
  }
  for (int j = 0; j < lb; ++j)
    n = ggm_step(k, tb + j, n, ((uint32_t)lane >> (lb - 1 - j)) & 1u);
  return n;
}

// Tile roots of a lane-DFS launch: thread (key, tile) walks the key's tb prefix
// levels to the tile's root. One key per block row (blockIdx.y), so the key is
// wave-uniform and its correction words come through scalar loads.
__global__ __launch_bounds__(256) void lane_tile_roots(LaneDfsParams p, uint4* __restrict__ seeds,
                                                       uint32_t* __restrict__ ts) {
  const uint64_t n_tiles = p.tile_end - p.tile_begin;
  const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n_tiles) return;
  const uint32_t key_i = blockIdx.y;
  const DevKey* __restrict__ k = p.keys + (p.key_idx ? p.key_idx[key_i] : key_i);
  const uint64_t tile = p.tile_begin + j;
  const int tb = (int)p.tb;
  Node n = key_root(k);
  for (int lvl = 0; lvl < tb; ++lvl) n = ggm_step(k, lvl, n, (uint32_t)(tile >> (tb - 1 - lvl)) & 1u);
  const size_t e = (size_t)key_i * n_tiles + j;
  seeds[e] = make_uint4(n.s0, n.s1, n.s2, n.s3);
  ts[e] = n.t;
}

// -------------------------------------------------------- lane-DFS scan ----

template <int C, bool kLateLoads, bool kGlobalStack, bool kRootTable>
__global__ __launch_bounds__(256) void scan_lane_dfs(LaneDfsParams p) {
  // kRootTable (with global stack levels only): tile roots come from the launch's table
  // instead of a walk; its own instance, so each instance has one way to a tile root.
  static_assert(!kRootTable || kGlobalStack, "the tile-root table goes with global levels");
  extern __shared__ uint4 lds[];
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t total = p.parts * p.n_keys;
  const uint32_t w = xcd_logical_wave(total, blockDim.x >> 6, wib);
  if (w - (uint32_t)wib >= total) return;  // the whole block is past the work
  const bool live = w < total;
  const uint32_t key_i = live ? w % p.n_keys : 0;
  const uint32_t part = live ? w / p.n_keys : 0;
  const DevKey* __restrict__ k = p.keys + (p.key_idx ? p.key_idx[key_i] : key_i);
  const int kl = (int)p.kl, lb = (int)p.lb, tb = (int)p.tb, d = (int)p.d;

  // Per-wave DFS stack: the right children waiting at depths 1 .. kl-1 of
  // the lane subtree (depth 0, the lane root, is never pushed). Seeds of
  // depths 1 .. gs sit in global memory (a wave's own [gs][64] region), the
  // deeper ones in LDS, [kl-1-gs][64]: depth-l slots are used 2^(l-1) times a
  // tile, so the few top ones cost nothing in global memory while their LDS
  // would limit the waves per CU. Each child's control byte is kept as the one
  // bit that matters, t == 1 (dpf.rs:221-239 and 136-141 test nothing else),
  // in bit `depth` of a per-lane register.
  // (kGlobalStack false: gslots is 0 and none of the global-slot code exists.)
  // The two parts are typed by address space, so a pop stays a ds_read or a
  // global load: through one generic pointer it would be a flat load, slower
  // for the LDS slots that nearly every pop reads.
  // With global levels, the wave's region of p.gstack also caches the
  // wave-uniform tile-prefix path (kLanePathU4 uint4s after the levels), which
  // the LDS-only instances keep in registers instead.
  const int gs = kGlobalStack ? (int)p.gslots : 0;
  lds_u32x4* st_seed = (lds_u32x4*)(lds + (size_t)wib * (kl - 1 - gs) * 64);  // slot s >= gs
  glob_u32x4* st_glob =
      kGlobalStack
          ? (glob_u32x4*)(p.gstack + (size_t)(live ? w : 0) * ((size_t)gs * 64 + kLanePathU4))
          : nullptr;
  glob_u32x4* path_seed = kGlobalStack ? st_glob + gs * 64 : nullptr;
  glob_u32* path_t = kGlobalStack ? (glob_u32*)(path_seed + 64) : nullptr;
  // This key's row of the launch's tile-root table.
  const glob_u32x4* root_seed = nullptr;
  const glob_u32* root_t = nullptr;
  if constexpr (kRootTable) {
    const size_t row = (size_t)key_i * (p.tile_end - p.tile_begin);
    root_seed = (const glob_u32x4*)p.root_seed + row;
    root_t = (const glob_u32*)p.root_t + row;
  }
  uint32_t st_tbits = 0;

  uint4 acc[C];
#pragma unroll
  for (int c = 0; c < C; c++) acc[c] = make_uint4(0, 0, 0, 0);

  // This wave's leaf-pair steps: global steps gs0 .. gs0 + n_steps - 1, step gs
  // being pair gs mod 2^(kl-1) of tile tile_begin + gs / 2^(kl-1).
  // (A key's steps number N / 2^(lb+1) <= 2^25, so a wave's counts fit 32 bits.)
  const uint32_t n_pairs = 1u << (kl - 1);
  const uint64_t total_steps = (p.tile_end - p.tile_begin) << (kl - 1);
  const uint32_t spw = (uint32_t)p.steps_per_wave;
  const uint64_t gs0 = (uint64_t)part * spw;
  uint32_t n_steps = 0;
  if (live && gs0 < total_steps)
    n_steps = total_steps - gs0 < spw ? (uint32_t)(total_steps - gs0) : spw;
  const bool lane_active = lane < (1 << lb);
  const int leaf_level = d - kl;

  auto push = [&](int slot, const Node& R) {  // the right child at depth slot + 1
    const u32x4 rv = {R.s0, R.s1, R.s2, R.s3};
    if (kGlobalStack && slot < gs)
      st_glob[slot * 64 + lane] = rv;
    else
      st_seed[(slot - gs) * 64 + lane] = rv;
    const uint32_t bit = 1u << (slot + 1);
    st_tbits = R.t == 1u ? (st_tbits | bit) : (st_tbits & ~bit);
  };

  LanePath path;  // (LDS-only instances)
  if constexpr (!kGlobalStack) {
    const Node r = key_root(k);
    path.s0 = r.s0; path.s1 = r.s1; path.s2 = r.s2; path.s3 = r.s3; path.t = r.t;
  }
  Node node;
  uint64_t tile_row0 = 0;  // the tile's first row; lane l's subtree starts 2^kl * l rows on
  // One iteration per leaf pair: pop the pending right child (if any), walk
  // down to the bottom node pushing right children, expand it into the two
  // leaves and fold the pair's rows into the accumulators. The rows are
  // loaded at the start of the step (kLateLoads false: in flight during the
  // expansions) or after the expansions (true: the rows stay out of the
  // registers the ChaCha rounds need, which with the global-stack levels and
  // the tile-root table leaves 32-byte rows at 60 VGPRs, eight waves per SIMD,
  // and the other waves on the SIMD run while a wave waits for its rows).
  // Loads are unconditional: a row past
  // n_rows (or of an idle lane) is read from a zero buffer instead.
  // The block's waves (consecutive keys, and parts) take the same number of
  // steps and meet at a barrier every 2^(kl-1) of them (every tile when runs
  // start at tile boundaries), so they stay in step and their identical row
  // reads hit in L2 instead of each fetching the rows again. Waves past the
  // work, and waves with shorter runs, take the same barriers.
  for (uint32_t s = 0; s < spw; ++s) {
    if (s && (s & (n_pairs - 1)) == 0) __syncthreads();
    if (s >= n_steps) continue;
    const uint64_t step = gs0 + s;  // (gs is the count of global stack levels)
    const uint64_t tile = p.tile_begin + (step >> (kl - 1));
    const uint32_t pair = (uint32_t)(step & (n_pairs - 1));
    int lvl;
    if (s == 0 || pair == 0) {  // the run's first step, or a new tile
      if constexpr (kRootTable) {
        {  // the tile's root from the launch's table, then the lane's lb levels
          const uint64_t j = tile - p.tile_begin;
          const u32x4 v = root_seed[j];
          Node n;
          n.s0 = v.x; n.s1 = v.y; n.s2 = v.z; n.s3 = v.w;
          n.t = root_t[j];
          for (int j = 0; j < lb; ++j)
            n = ggm_step(k, tb + j, n, ((uint32_t)lane >> (lb - 1 - j)) & 1u);
          node = n;
        }
      } else if constexpr (kGlobalStack) {
        node = tile_lane_root_cached(k, path_seed, path_t, lane, tile, tile - 1, s == 0, tb, lb);
      } else
        node = tile_lane_root(k, path, lane, tile, tile - 1, s == 0, tb, lb);
      tile_row0 = tile << (kl + lb);
      lvl = 0;
      if (pair != 0) {
        // The run starts inside a tile: walk from the lane root down to leaf
        // pair `pair`, keeping the right children that lie ahead of it, as the
        // depth-first walk from pair 0 would have left them.
        for (int l = 0; l < kl - 1; ++l) {
          Node L, R;
          expand_node(k, leaf_level + l, node, L, R);
          if ((pair >> (kl - 2 - l)) & 1u) {
            node = R;
          } else {
            push(l, R);
            node = L;
          }
        }
        lvl = kl - 1;
      }
    } else {  // the right child at depth lvl >= 1 sits in slot lvl - 1
      lvl = kl - 1 - __builtin_ctz(pair);
      u32x4 sv;
      if (kGlobalStack && lvl - 1 < gs)
        sv = st_glob[(lvl - 1) * 64 + lane];
      else
        sv = st_seed[(lvl - 1 - gs) * 64 + lane];
      node.s0 = sv.x; node.s1 = sv.y; node.s2 = sv.z; node.s3 = sv.w;
      node.t = (st_tbits >> lvl) & 1u;
    }
    uint4 r0[C], r1[C];
    auto load_rows = [&]() {
      // (the lane bits are added here rather than kept in a register pair)
      const uint64_t x = (tile_row0 | ((uint64_t)__lane_id() << kl)) + 2 * (uint64_t)pair;
      const uint4* src0 = (lane_active && x < p.n_rows) ? p.rows + x * C : p.zero_rows;
      const uint4* src1 =
          (lane_active && x + 1 < p.n_rows) ? p.rows + (x + 1) * C : p.zero_rows;
#pragma unroll
      for (int c = 0; c < C; c++) {
        r0[c] = src0[c];
        r1[c] = src1[c];
      }
    };
    if constexpr (!kLateLoads) load_rows();
    for (; lvl < kl - 1; ++lvl) {
      Node L, R;
      expand_node(k, leaf_level + lvl, node, L, R);
      push(lvl, R);  // depth lvl + 1
      node = L;
    }
    Node L, R;
    expand_node(k, leaf_level + kl - 1, node, L, R);
    if constexpr (kLateLoads) load_rows();
    const uint4 m0 = opaque4(leaf_mask(k, L)), m1 = opaque4(leaf_mask(k, R));
#pragma unroll
    for (int c = 0; c < C; c++) {  // one v_bitop3_b32 per word and row
      xor_masked3(acc[c], r0[c], m0);
      xor_masked3(acc[c], r1[c], m1);
    }
  }

#pragma unroll
  for (int c = 0; c < C; c++) {
    const uint4 v = wave_xor4(acc[c]);
    if (live && lane == 0) p.partial[((size_t)part * p.n_keys + key_i) * C + c] = v;
  }
}

// ------------------------------------------------------ point-eval scan ----
// One lane per row; the row index is rows_idx[e] when given (delta entries),
// else e itself. Per-wave XOR partials per key.
__global__ __launch_bounds__(256) void scan_point(PointParams p) {
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t w = blockIdx.x * (blockDim.x >> 6) + wib;  // entry wave
  if ((uint64_t)w * 64 >= p.n_entries) return;
  const uint32_t key_i = blockIdx.y;
  const DevKey* __restrict__ k = p.keys + (p.key_idx ? p.key_idx[key_i] : key_i);
  const uint64_t e = (uint64_t)w * 64 + lane;
  const bool valid = e < p.n_entries;
  const uint64_t x = valid ? (p.row_idx ? p.row_idx[e] : p.first_row + e) : 0;
  const uint4 m = valid ? point_eval(k, x) : make_uint4(0, 0, 0, 0);
  const uint4* src = p.rows + (p.row_idx ? e : x) * p.chunks;
  for (uint32_t c = 0; c < p.chunks; c++) {
    uint4 a = make_uint4(0, 0, 0, 0);
    if (valid) xor_masked(a, src[c], m);
    a = wave_xor4(a);
    if (lane == 0) p.partial[((size_t)w * p.n_keys + key_i) * p.chunks + c] = a;
  }
}

// ---------------------------------------------------- suffix-walk baseline --
template <int C>
__global__ __launch_bounds__(256) void scan_suffix(SuffixParams p) {
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t t = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wib;
  if (t >= p.n_tiles) return;
  const uint32_t key_i = blockIdx.y;
  const DevKey* __restrict__ k = p.keys + (p.key_idx ? p.key_idx[key_i] : key_i);
  const int d = (int)p.d, s = (int)p.s;
  const uint64_t tile = p.tile_begin + t;
  const uint64_t start = tile << s;
  Node root = key_root(k);
  for (int lvl = 0; lvl < d - s; lvl++)
    root = ggm_step(k, lvl, root, (uint32_t)(start >> (d - 1 - lvl)) & 1u);
  uint4 acc[C];
#pragma unroll
  for (int c = 0; c < C; c++) acc[c] = make_uint4(0, 0, 0, 0);
  const uint32_t tile_rows = 1u << s;
  for (uint32_t it = 0; it < tile_rows; it += 64) {
    const uint32_t off = it + lane;
    const uint64_t x = start + off;
    const bool valid = off < tile_rows && x < p.n_rows;
    Node n = root;
    for (int j = 0; j < s; j++)
      n = ggm_step(k, d - s + j, n, (uint32_t)(x >> (s - 1 - j)) & 1u);
    const uint4 m = leaf_mask(k, n);
    if (valid) {
      const uint4* src = p.rows + x * C;
#pragma unroll
      for (int c = 0; c < C; c++) xor_masked(acc[c], src[c], m);
    }
  }
#pragma unroll
  for (int c = 0; c < C; c++) {
    const uint4 v = wave_xor4(acc[c]);
    if (lane == 0) p.partial[((size_t)t * p.n_keys + key_i) * C + c] = v;
  }
}

// ------------------------------------------------------- stream (wide) ----
// G keys per wave, rows of C = 64*S chunks or fewer. Every leaf-pair step of
// the lane DFS leaves 2 masks per lane per key in LDS; the wave then reads
// those 128 rows once each, coalesced (lane = chunk), and folds them into G
// accumulators. Accumulator slot j of lane l holds chunk l + 64*j.
// Narrow rows (C <= 64 / PK, S = 1) are packed PK to a load: lane l reads
// chunk l % (64 / PK) of row l / (64 / PK) of each group of PK rows, so a
// load instruction, an LDS mask read and a fold carry PK rows instead of
// one. The PK lanes that hold one chunk are XOR-combined at the end.
template <int S, int G, int PK>
__global__ __launch_bounds__(256) void scan_stream(StreamParams p) {
  static_assert(PK == 1 || S == 1, "rows are packed only when one load covers a row");
  constexpr int P = 64 / PK;  // lanes per row
  extern __shared__ uint4 lds[];
  const int lane = threadIdx.x & 63;
  const uint32_t sub = PK > 1 ? (uint32_t)lane / P : 0;  // this lane's row within a load
  const uint32_t lane_chunk = PK > 1 ? (uint32_t)lane % P : (uint32_t)lane;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t n_groups = (p.n_keys + G - 1) / G;
  const uint32_t w = xcd_logical_wave(p.parts * n_groups, blockDim.x >> 6, wib);
  if (w >= p.parts * n_groups) return;
  const uint32_t grp = w % n_groups;
  const uint32_t part = w / n_groups;
  const uint32_t key0 = grp * G;
  const int kl = (int)p.kl, lb = (int)p.lb, tb = (int)p.tb, d = (int)p.d;
  const uint32_t C = p.chunks;

  // LDS per wave: G stacks of seeds at depths 1 .. kl-1, [kl-1][64] each,
  // then masks [G][128]. Each key's pushed control bytes are kept as the bit
  // t == 1 in a register (bit `depth`), as in scan_lane_dfs.
  const size_t per_wave = (size_t)G * (kl - 1) * 64 + (size_t)G * 128;
  uint4* base_lds = lds + (size_t)wib * per_wave;
  uint4* masks = base_lds + (size_t)G * (kl - 1) * 64;
  uint32_t st_tbits[G];
#pragma unroll
  for (int g = 0; g < G; g++) st_tbits[g] = 0;

  const DevKey* kp[G];
  bool key_live[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    key_live[g] = key0 + g < p.n_keys;
    const uint32_t ki = key_live[g] ? key0 + g : key0;
    kp[g] = p.keys + (p.key_idx ? p.key_idx[ki] : ki);
  }

  uint4 acc[G][S];
#pragma unroll
  for (int g = 0; g < G; g++)
#pragma unroll
    for (int j = 0; j < S; j++) acc[g][j] = make_uint4(0, 0, 0, 0);

  const uint64_t t_begin = p.tile_begin + (uint64_t)part * p.tiles_per_wave;
  uint64_t t_end = t_begin + p.tiles_per_wave;
  if (t_end > p.tile_end) t_end = p.tile_end;
  const uint32_t n_pairs = 1u << (kl - 1);
  const int leaf_level = d - kl;
  const int lanes_live = 1 << lb;

  LanePath path[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    const Node r = key_root(kp[g]);
    path[g].s0 = r.s0; path[g].s1 = r.s1; path[g].s2 = r.s2;
    path[g].s3 = r.s3; path[g].t = r.t;
  }

  // Rows of one leaf-pair step: 2 per live source lane s, tile_base + s*2^kl +
  // 2*pair + {0,1}; row r of the step uses mask slot r. They are read in
  // batches of U rows (IU load instructions: 8 KiB of 1 KiB rows), NBUF = 2
  // batches in flight: the first two batches of a step are issued before
  // that step's PRG work, so loads overlap the ChaCha rounds. Three or four
  // batches keep more of a step in flight but hold their registers through
  // the ChaCha rounds: with four, one wave fits per SIMD instead of two, and
  // 144-byte rows with one key slow from 9.1 to 11.5 ms (three: 8.9 ms for
  // one key, 15.9 against 14.4 ms for two).
  // Loads are branch-free: a row's base address is wave-uniform (a row past
  // the step or the matrix reads the zero buffer instead), and each lane's
  // chunk offset is fixed for the whole kernel. A lane past the row's last
  // chunk re-reads chunk 0; its accumulator is never written out.
  constexpr int IU = 8 / S;   // load instructions per batch
  constexpr int U = IU * PK;  // rows per batch
  constexpr int NBUF = 2;
  const uint32_t step_rows = 2u * (uint32_t)lanes_live;
  uint32_t lane_off[S];
#pragma unroll
  for (int j = 0; j < S; j++) {
    const uint32_t c = lane_chunk + 64 * j;
    lane_off[j] = c < C ? c : 0;
  }
  // In a tile that lies wholly inside the matrix (every tile but a ragged
  // last one) a batch's rows sit at fixed distances from its first row: row
  // r0 + q (r0 even) is (q >> 1) lane subtrees of 2^kl rows and (q & 1) rows
  // further on. This lane reads row q = i*PK + sub with load i. Those byte
  // offsets, with this lane's chunk, are computed once, so a batch costs one
  // uniform base address and IU x S plain loads.
  uint32_t row_off[IU][S];
#pragma unroll
  for (int i = 0; i < IU; i++)
#pragma unroll
    for (int j = 0; j < S; j++) {
      const uint32_t q = (uint32_t)i * PK + sub;
      row_off[i][j] =
          (uint32_t)(((((uint64_t)(q >> 1) << kl) + (q & 1)) * C + lane_off[j]) * 16);
    }
  uint4 ring[NBUF][IU][S];
  auto load_batch = [&](uint64_t tile_base, uint32_t pair, uint32_t r0, uint4 (&v)[IU][S],
                        bool whole_tile) {
    if (whole_tile) {
      const char* base = reinterpret_cast<const char*>(
          p.rows + (tile_base + ((uint64_t)(r0 >> 1) << kl) + 2 * (uint64_t)pair) * C);
#pragma unroll
      for (int i = 0; i < IU; i++)
#pragma unroll
        for (int j = 0; j < S; j++)
          v[i][j] = load_nt(reinterpret_cast<const uint4*>(base + row_off[i][j]));
      return;
    }
#pragma unroll
    for (int i = 0; i < IU; i++) {
      const uint32_t r = r0 + (uint32_t)i * PK + sub;
      const uint64_t x = tile_base + ((uint64_t)(r >> 1) << kl) + 2 * (uint64_t)pair + (r & 1);
      const uint4* src = (r < step_rows && x < p.n_rows) ? p.rows + x * C : p.zero_rows;
#pragma unroll
      for (int j = 0; j < S; j++) v[i][j] = load_nt(src + lane_off[j]);
    }
  };
  // Slot indices stay below 128 (r0 + q < max(step_rows, U) <= 128), so they
  // need no wrap.
  auto fold_batch = [&](uint32_t r0, const uint4 (&v)[IU][S]) {
#pragma unroll
    for (int i = 0; i < IU; i++) {
#pragma unroll
      for (int g = 0; g < G; g++) {
        const uint4 m = masks[g * 128 + r0 + i * PK + sub];
#pragma unroll
        for (int j = 0; j < S; j++) xor_masked3(acc[g][j], v[i][j], m);
      }
    }
  };
  auto prefetch_step = [&](uint64_t tile_base, uint32_t pair, bool whole_tile) {
#pragma unroll
    for (int k = 0; k < NBUF; k++)
      if (k == 0 || (uint32_t)k * U < step_rows)
        load_batch(tile_base, pair, (uint32_t)k * U, ring[k], whole_tile);
  };

  for (uint64_t tile = t_begin; tile < t_end; ++tile) {
    const uint64_t tile_base = tile << (kl + lb);
    // Every row a step of this tile reads exists, and batches never run past
    // the step (step_rows is a power of two >= U).
    const bool whole_tile =
        step_rows >= (uint32_t)U && tile_base + (1ull << (kl + lb)) <= p.n_rows;
    prefetch_step(tile_base, 0, whole_tile);
    Node node[G];
#pragma unroll
    for (int g = 0; g < G; g++)
      node[g] = tile_lane_root(kp[g], path[g], lane, tile, tile - 1, tile == t_begin, tb, lb);
    int lvl = 0;
    uint32_t pair = 0;
    for (;;) {
#pragma unroll
      for (int g = 0; g < G; g++) {
        uint4* st_seed = base_lds + (size_t)g * (kl - 1) * 64;
        Node L, R;
        expand_node<kStreamPrgVariant, false>(kp[g], leaf_level + lvl, node[g], L, R);
        if (lvl == kl - 1) {
          masks[g * 128 + 2 * lane] = leaf_mask<false>(kp[g], L);
          masks[g * 128 + 2 * lane + 1] = leaf_mask<false>(kp[g], R);
        } else {
          st_seed[lvl * 64 + lane] = make_uint4(R.s0, R.s1, R.s2, R.s3);  // depth lvl + 1
          const uint32_t bit = 1u << (lvl + 1);
          st_tbits[g] = R.t == 1u ? (st_tbits[g] | bit) : (st_tbits[g] & ~bit);
          node[g] = L;
        }
      }
      if (lvl < kl - 1) {
        ++lvl;
        continue;
      }
      // ring[k] already holds batch k of this step; batch b + NBUF takes the
      // slot that batch b frees.
      for (uint32_t r0 = 0; r0 < step_rows; r0 += NBUF * U) {
#pragma unroll
        for (int k = 0; k < NBUF; k++) {
          const uint32_t rk = r0 + (uint32_t)k * U;
          if (k > 0 && rk >= step_rows) break;
          fold_batch(rk, ring[k]);
          if (rk + NBUF * U < step_rows)
            load_batch(tile_base, pair, rk + NBUF * U, ring[k], whole_tile);
        }
      }
      if (++pair == n_pairs) break;
      prefetch_step(tile_base, pair, whole_tile);
      lvl = kl - 1 - __builtin_ctz(pair);
#pragma unroll
      for (int g = 0; g < G; g++) {  // the right child at depth lvl >= 1: slot lvl - 1
        const uint4* st_seed = base_lds + (size_t)g * (kl - 1) * 64;
        const uint4 sv = st_seed[(lvl - 1) * 64 + lane];
        node[g].s0 = sv.x; node[g].s1 = sv.y; node[g].s2 = sv.z; node[g].s3 = sv.w;
        node[g].t = (st_tbits[g] >> lvl) & 1u;
      }
    }
  }

#pragma unroll
  for (int g = 0; g < G; g++) {
    if (PK > 1) acc[g][0] = xor_lanes_from(acc[g][0], P);
    if (!key_live[g]) continue;
#pragma unroll
    for (int j = 0; j < S; j++) {
      const uint32_t c = lane_chunk + 64 * j;
      if (sub == 0 && c < C) p.partial[((size_t)part * p.n_keys + key0 + g) * C + c] = acc[g][j];
    }
  }
}

// ------------------------------------------- split producer/consumer scan --
// Wide rows, several keys per pass. A block holds NP producer/consumer pairs
// of waves: waves 0 .. DW*NP-1 run the keys' lane DFS as scan_stream does (DW
// per pair, each with G / DW of its keys) and publish each leaf-pair step's
// 2 x 64 masks per key into a ring of D mask buffers of their pair; the last
// NP waves only stream their pair's rows (NB batches of U rows in flight,
// including the next step's first batches while they wait) and fold them. The
// waves of a pair hand steps over through counters in LDS, steps published
// (one per DFS wave) and steps folded: a DFS
// wave runs up to D - 1 steps ahead and only waits when the ring is full, so
// the steps that descend a whole subtree (kl expansions per key, against one
// or two for most steps) no longer stall the streaming wave as a per-step
// block barrier did. Pairs never wait for each other. Every wait is bounded
// (kSplitSpinLimit polls); a wait that runs out sets *p.fault, which the host
// reports as an error, and the wave stops, so the kernel always ends.
// With NP = 4 the waves of one block land one producer and one consumer on
// each SIMD, where single-pair blocks can stack two DFS waves on one SIMD.
// Needs all 64 lanes live (lb = 6), so a step is always 128 rows.
constexpr int kSplitNB = 4;
constexpr uint32_t kSplitSpinLimit = 1u << 22;  // ~0.3 s of polls; a step takes microseconds
// Ring turns (NB batches) a streaming wave may run ahead of a partner pair reading the same
// rows for other keys.
constexpr uint32_t kSplitPace = 2;

__device__ __forceinline__ uint32_t lds_load_acquire(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_store_release(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waits until the counter has reached `want` (both count steps, so the
// difference stays small and the signed compare is wrap-safe). False when the
// wait ran out.
__device__ __forceinline__ bool wait_count(const uint32_t* ctr, uint32_t want) {
  for (uint32_t spin = 0; spin < kSplitSpinLimit; ++spin) {
    if ((int32_t)(lds_load_acquire(ctr) - want) >= 0) return true;
    __builtin_amdgcn_s_sleep(1);
  }
  return false;
}

// Stack slots of a key that the split kernel's DFS wave holds in registers:
// the top one (depth 1, used once per tile) and the R - 1 bottom ones (depths
// kl - 1, kl - 2, ..., pushed and popped within a few steps). Groups of four
// keys hold two more bottom levels in registers, which leaves their four pairs
// of a block the LDS for a ring of four mask steps at kl = 7, or three at kl = 9.
__host__ __device__ constexpr int split_reg_slots(int G) { return G >= 4 ? 5 : 3; }

// Stack levels of a depth-kl lane DFS kept in LDS by the split kernel: the
// kl - 1 levels minus the register slots.
__host__ __device__ __forceinline__ uint32_t split_lds_levels(int G, uint32_t kl) {
  const uint32_t r = (uint32_t)split_reg_slots(G);
  return kl > r + 1 ? kl - r - 1 : 0;
}

// LDS of one pair, in uint4 units: G stacks of split_lds_levels x 64 seeds,
// D mask buffers of G x 128 masks, and one uint4 holding the counters (steps
// published by the first DFS wave, steps folded, steps published by the second).
__host__ __device__ __forceinline__ size_t split_pair_u4(int G, uint32_t kl, uint32_t D) {
  return (size_t)G * split_lds_levels(G, kl) * 64 + (size_t)D * G * 128 + 1;
}

// DW DFS waves per pair (1 or 2): with two, each makes the masks of G / DW of
// the pair's keys, so the keys' ChaCha work runs on two waves of a SIMD, and
// the streaming wave waits for both waves' counters. Blocks of DW = 2 hold
// 3 NP waves and keep two row batches in flight instead of four, which leaves
// the registers for three waves per SIMD.
template <int S, int G, int NP, int DW>
__global__ __launch_bounds__(64 * (DW + 1) * NP) void scan_split(StreamParams p) {
  static_assert(DW == 1 || DW == 2, "one or two DFS waves per pair");
  static_assert(G % DW == 0, "a pair's keys split evenly over its DFS waves");
  constexpr int GW = G / DW;  // keys per DFS wave
  // Two DFS waves share a SIMD, and the stage-spaced node routine pays there (4 keys of
  // 78,643,200 x 1 KiB, one process: 12.13 / 12.29 against 12.43 / 12.54 ms); one DFS
  // wave keeps the unspaced one.
  constexpr int kVar = DW == 2 ? kPrgSpaced : kSplitPrgVariant;
  extern __shared__ uint4 lds[];
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool producer = wib < NP * DW;
  const int pi = producer ? wib / DW : wib - NP * DW;  // this wave's pair in the block
  const int sub = producer ? wib % DW : 0;            // a DFS wave's share of the keys
  const uint32_t n_groups = (p.n_keys + G - 1) / G;
  const uint32_t halves = p.col_splits;
  const uint32_t total = p.parts * n_groups * halves;  // pairs with work
  const uint32_t per = (gridDim.x + kXcds - 1) / kXcds;
  const uint32_t blk = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;  // XCD-aware
  if ((uint64_t)blk * NP >= total) return;  // the whole block
  // Pairs that split rows into column halves (rows of 2 chunks per lane in each half, blocks
  // of 2 or 4 pairs) share their DFS work: pair 2j + 1 of a block takes the other half of
  // the rows of pair 2j's part and group, and its streaming wave folds them with pair 2j's
  // masks, so one DFS wave makes the masks for both halves (the other DFS wave ends at
  // once), and pair 2j's DFS wave waits for both streaming waves before it reuses a buffer
  // (16,817,159 x 4 KiB, 24 / 96 keys: 40.9 / 150.6 ms a pass and 1.2-1.3x / 2.6-3.2x the
  // matrix from HBM, against 50.8 / 169.3 ms and 2.24x / 4.1x with four groups of a half
  // per block and the masks made twice).
  constexpr bool kHalfShare = NP > 1 && DW == 1 && S == 2;
  const bool share = kHalfShare && halves == 2;
  // Pair order: column half fastest, then group, then part when halves share masks; else
  // group fastest, then column half, then part. Either way the pairs of a block read the
  // same rows, and in passes of several groups the block's pairs of one half are NP / 2
  // (or NP) groups reading the same chunks of them, so that one HBM read serves them all.
  const uint32_t pid = blk * NP + (uint32_t)pi;
  const bool live = pid < total;
  // Column half of pair index q (its group: (q >> 1) % n_groups when halves share masks,
  // else q % n_groups).
  auto half_of = [&](uint32_t q) -> uint32_t {
    return share ? (q & 1u) : (q / n_groups) % halves;
  };
  uint32_t grp, half;
  if constexpr (kHalfShare) {
    grp = live ? (share ? (pid >> 1) % n_groups : pid % n_groups) : 0;
    half = live ? half_of(pid) : 0;
  } else {
    // (Other instances: column half fastest, then group.)
    half = live ? pid % halves : 0;
    grp = live ? (pid / halves) % n_groups : 0;
  }
  const uint32_t part = live ? pid / (halves * n_groups) : 0;
  const uint32_t key0 = grp * G;
  const int kl = (int)p.kl, lb = (int)p.lb, tb = (int)p.tb, d = (int)p.d;
  const uint32_t C = p.chunks;  // row stride in chunks
  const uint32_t col0 = half * 64 * S;  // this pair's chunks: [col0, col_end)
  const uint32_t col_end = halves == 1 ? C : (col0 + 64 * S < C ? col0 + 64 * S : C);
  // This pair's leaf-pair steps: global steps gs0 .. gs0 + n_steps - 1, where global step
  // gs is pair gs mod 2^(kl-1) of tile tile_begin + gs / 2^(kl-1).
  const uint32_t n_pairs = 1u << (kl - 1);
  const uint64_t total_steps = (p.tile_end - p.tile_begin) << (kl - 1);
  const uint64_t gs0 = (uint64_t)part * p.steps_per_wave;
  uint64_t n_steps = 0;
  if (live && gs0 < total_steps)
    n_steps = total_steps - gs0 < p.steps_per_wave ? total_steps - gs0 : p.steps_per_wave;
  const uint32_t D = p.mask_depth;
  // Right children wait at depths 1 .. kl-1 (depth 0 is the lane root,
  // never pushed): a key's stack has kl - 1 slots, slot l holding depth l + 1.
  // Slot 0 and the R - 1 bottom slots kl - 2, kl - 3, ... sit in the DFS
  // wave's registers, the others in LDS ([G][kl - 1 - R][64] seeds, slot l at
  // level l - 1); the control bytes are kept as the bit t == 1 in a register
  // per key.
  constexpr int R = split_reg_slots(G);
  const int slots = kl - 1;
  const int lds_slots = (int)split_lds_levels(G, p.kl);
  uint4* pair_lds = lds + (size_t)pi * split_pair_u4(G, p.kl, D);
  uint4* stacks = pair_lds;                                  // [G][lds_slots][64]
  uint4* masks = pair_lds + (size_t)G * lds_slots * 64;      // [D][G][128]
  uint32_t* ctr = reinterpret_cast<uint32_t*>(masks + (size_t)D * G * 128);
  // Pair q's counters: [0] published, [1] folded, [2] published by a second DFS wave,
  // [3] ring turns of the streaming wave.
  auto pair_ctr = [&](int q) -> uint32_t* {
    uint4* ql = lds + (size_t)q * split_pair_u4(G, p.kl, D);
    return reinterpret_cast<uint32_t*>(ql + (size_t)G * lds_slots * 64 + (size_t)D * G * 128);
  };
  uint32_t* published = ctr + 2 * sub;  // steps whose masks (this DFS wave's keys) are in the ring
  uint32_t* folded = ctr + 1;  // steps whose masks the streaming wave is done with
  uint32_t* turns = ctr + 3;  // ring turns the streaming wave has started
  if (producer && lane == 0) {
    *published = 0;
    if (sub == 0) {
      *folded = 0;
      *turns = 0;
    }
  }
  __syncthreads();  // the only block barrier: counters are set before any wave reads them

  if constexpr (kHalfShare) {
    if (producer && share && half == 1) return;  // pair 2j's DFS wave makes these masks
  }
  if (producer) {
    // This wave's keys are the pair's keys sub * GW .. sub * GW + GW - 1; arrays
    // below are indexed by g < GW, the pair's LDS stacks and masks by gk = sub * GW + g.
    const int gk0 = sub * GW;
    const DevKey* kp[GW];
#pragma unroll
    for (int g = 0; g < GW; g++) {
      const uint32_t ki = key0 + gk0 + g < p.n_keys ? key0 + gk0 + g : key0;
      kp[g] = p.keys + (p.key_idx ? p.key_idx[ki] : ki);
    }
    LanePath path[GW];
#pragma unroll
    for (int g = 0; g < GW; g++) {
      const Node r = key_root(kp[g]);
      path[g].s0 = r.s0; path[g].s1 = r.s1; path[g].s2 = r.s2;
      path[g].s3 = r.s3; path[g].t = r.t;
    }
    const int leaf_level = d - kl;
    Node node[GW];
    uint32_t st_tbits[GW];
    // Register slots as plain words: [0] slot 0, [r] slot kl - 1 - r for r >= 1 (for
    // small kl they may name the same slot; each holds what was last pushed there).
    uint32_t st_reg[GW][R][4];
#pragma unroll
    for (int g = 0; g < GW; g++) st_tbits[g] = 0;
    // Seed of stack slot `slot` (uniform) for key g. The register slots are
    // written and read as values (selects), never through a pointer, so they
    // stay in VGPRs.
    auto put_seed = [&](int g, int slot, const Node& n) __attribute__((always_inline)) {
      const uint32_t v[4] = {n.s0, n.s1, n.s2, n.s3};
      bool any = false;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const bool in_reg = slot == (r == 0 ? 0 : slots - r);
        any |= in_reg;
#pragma unroll
        for (int i = 0; i < 4; i++) st_reg[g][r][i] = in_reg ? v[i] : st_reg[g][r][i];
      }
      if (!any)
        stacks[((size_t)(gk0 + g) * lds_slots + (slot - 1)) * 64 + lane] =
            make_uint4(v[0], v[1], v[2], v[3]);
    };
    auto get_seed = [&](int g, int slot) __attribute__((always_inline)) -> uint4 {
      uint32_t a[R][4];
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          a[r][i] = st_reg[g][r][i];
          // Every word is read as a value: folding the selects below into a load
          // through a selected address would keep the slots in scratch memory.
          asm("" : "+v"(a[r][i]));
        }
      // The register slot holding `slot`, counting from the last one (R - 1 when none).
      uint32_t w[4] = {a[R - 1][0], a[R - 1][1], a[R - 1][2], a[R - 1][3]};
      bool any = slot == slots - (R - 1);
#pragma unroll
      for (int r = R - 2; r >= 0; r--) {
        const bool hit = slot == (r == 0 ? 0 : slots - r);
        any |= hit;
#pragma unroll
        for (int i = 0; i < 4; i++) w[i] = hit ? a[r][i] : w[i];
      }
      uint4 v = make_uint4(w[0], w[1], w[2], w[3]);
      if (!any) v = stacks[((size_t)(gk0 + g) * lds_slots + (slot - 1)) * 64 + lane];
      return v;
    };
    // The next tile's lane roots, walked one tree level per step while the
    // current tile runs (levels pre_lvl .. tb + lb - 1 still to go), so that a
    // tile's first step only descends its kl levels. Walking the changed
    // tile-prefix levels and the lb lane levels (about 8 node expansions per
    // key) all in that step held the streaming wave up once per tile.
    Node nroot[GW];
    int pre_lvl = -1;  // -1: no walk under way
    const int root_lvl = tb + lb;
    auto root_level = [&](uint64_t ntile) {  // one level of the walk, every key
#pragma unroll
      for (int g = 0; g < GW; g++) {
        const uint32_t bit = pre_lvl < tb ? (uint32_t)(ntile >> (tb - 1 - pre_lvl)) & 1u
                                          : ((uint32_t)lane >> (root_lvl - 1 - pre_lvl)) & 1u;
        Node a, b;
        expand_node<kVar, false>(kp[g], pre_lvl, nroot[g], a, b);
        nroot[g] = bit ? b : a;
        if (pre_lvl < tb) path[g].put(lane, pre_lvl + 1, nroot[g]);
      }
      ++pre_lvl;
    };
    for (uint64_t s = 0; s < n_steps; ++s) {
      // Buffer s % D is free once the streaming wave has folded step s - D (both halves'
      // streaming waves, when they share the masks).
      if (s >= D && !wait_count(folded, (uint32_t)(s - D + 1))) {
        atomicOr(p.fault, 1u);
        return;
      }
      if constexpr (kHalfShare) {
        if (s >= D && share && !wait_count(pair_ctr(pi + 1) + 1, (uint32_t)(s - D + 1))) {
          atomicOr(p.fault, 1u);
          return;
        }
      }
      const uint64_t gs = gs0 + s;
      const uint64_t tile = p.tile_begin + (gs >> (kl - 1));
      const uint32_t pair = (uint32_t)(gs & (n_pairs - 1));
      const bool from_root = pair == 0 || s == 0;  // a new tile, or the run's first step
      if (from_root) {
        if (s != 0 && pre_lvl >= 0) {
          while (pre_lvl < root_lvl) root_level(tile);  // finish the walk
#pragma unroll
          for (int g = 0; g < GW; g++) node[g] = nroot[g];
        } else {
#pragma unroll
          for (int g = 0; g < GW; g++)
            node[g] = tile_lane_root(kp[g], path[g], lane, tile, tile - 1, s == 0, tb, lb);
        }
        pre_lvl = -1;
      }
      const bool seek = s == 0 && pair != 0;
      if (seek) {
        // The run starts inside a tile: walk each key from the lane root down to leaf
        // pair `pair`, keeping the right children that lie ahead of it, as the
        // depth-first walk from pair 0 would have left them.
#pragma unroll
        for (int g = 0; g < GW; g++) {
          for (int l = 0; l < kl - 1; ++l) {
            Node L, R;
            expand_node<kVar, false>(kp[g], leaf_level + l, node[g], L, R);
            if ((pair >> (kl - 2 - l)) & 1u) {
              node[g] = R;
            } else {
              put_seed(g, l, R);
              const uint32_t bit = 1u << (l + 1);
              st_tbits[g] = R.t == 1u ? (st_tbits[g] | bit) : (st_tbits[g] & ~bit);
              node[g] = L;
            }
          }
        }
      }
      uint4* mbuf = masks + (size_t)(s % D) * G * 128;
      const int lvl = seek ? kl - 1 : from_root ? 0 : kl - 1 - __builtin_ctz(pair);
#pragma unroll
      for (int g = 0; g < GW; g++) {
        if (!from_root) {  // the right child at depth lvl >= 1 sits in slot lvl - 1
          const uint4 sv = get_seed(g, lvl - 1);
          node[g].s0 = sv.x; node[g].s1 = sv.y; node[g].s2 = sv.z; node[g].s3 = sv.w;
          node[g].t = (st_tbits[g] >> lvl) & 1u;
        }
        Node L, R;
        for (int l = lvl; l < kl - 1; ++l) {
          expand_node<kVar, false>(kp[g], leaf_level + l, node[g], L, R);
          put_seed(g, l, R);  // depth l + 1
          const uint32_t bit = 1u << (l + 1);
          st_tbits[g] = R.t == 1u ? (st_tbits[g] | bit) : (st_tbits[g] & ~bit);
          node[g] = L;
        }
        expand_node<kVar, false>(kp[g], leaf_level + kl - 1, node[g], L, R);
        mbuf[(gk0 + g) * 128 + 2 * lane] = leaf_mask<false>(kp[g], L);
        mbuf[(gk0 + g) * 128 + 2 * lane + 1] = leaf_mask<false>(kp[g], R);
      }
      lds_store_release(published, (uint32_t)(s + 1));  // after this wave's mask stores
      // One level of the next tile's lane-root walk, if this run reaches that tile.
      if (((tile + 1 - p.tile_begin) << (kl - 1)) < gs0 + n_steps) {
        if (pre_lvl < 0) {  // start at the first level where the two tiles' paths part
          const uint64_t diff = (tile + 1) ^ tile;
          pre_lvl = tb - (64 - __builtin_clzll(diff));
#pragma unroll
          for (int g = 0; g < GW; g++) nroot[g] = path[g].get(pre_lvl);
        }
        if (pre_lvl < root_lvl) root_level(tile + 1);
      }
    }
    return;
  }

  // Consumer: rows of step s are tile_base + j*2^kl + 2*pair + {0,1} for
  // source lane j < 64; row r of the step uses mask slot r.
  constexpr int U = 8 / S;
  constexpr int NB = DW == 2 ? 2 : kSplitNB;
  const uint32_t n_batches = 128 / U;  // a multiple of NB
  uint32_t lane_off[S];
#pragma unroll
  for (int j = 0; j < S; j++) {
    const uint32_t c = col0 + lane + 64 * j;
    lane_off[j] = c < col_end ? c : 0;
  }
  uint32_t row_off[U][S];
#pragma unroll
  for (int u = 0; u < U; u++)
#pragma unroll
    for (int j = 0; j < S; j++)
      row_off[u][j] =
          (uint32_t)(((((uint64_t)(u >> 1) << kl) + (u & 1)) * C + lane_off[j]) * 16);
  uint4 acc[G][S];
#pragma unroll
  for (int g = 0; g < G; g++)
#pragma unroll
    for (int j = 0; j < S; j++) acc[g][j] = make_uint4(0, 0, 0, 0);
  // Partners: the block's other pairs of this part and column half (other key groups),
  // which read the same rows. Their streaming waves keep within kSplitPace ring turns of
  // each other, so that one pair's row reads still sit in L2 when the others ask for the
  // same rows, and a pass of several groups reads the matrix from HBM about once (without
  // the pacing, the pairs of a block drift apart: 1.55x the matrix at 8 keys over 2 KiB
  // rows). The pair that is furthest behind never waits, so the pacing cannot deadlock.
  // Only rows of 2 chunks per lane are paced (2 KiB rows, and the column halves of 4 KiB
  // rows that passes of 4 or more keys use): a pass of several groups over 1 KiB rows is
  // bound by its DFS waves, and tying the pairs together there cost 9-15 % (78,643,200 x
  // 1 KiB, 8 keys: 26.3-27.8 against 24.0 ms); whole 4 KiB rows run one group per pass
  // unless column halves are turned off. Pairs with two DFS waves run in passes of one
  // group.
  constexpr bool kPaced = NP > 1 && DW == 1 && S == 2;
  uint32_t partners = 0;
  if constexpr (kPaced) {
#pragma unroll
    for (int q = 0; q < NP; q++) {
      const uint32_t pq = blk * NP + (uint32_t)q;
      if (q != pi && live && pq < total && pq / (halves * n_groups) == part && half_of(pq) == half)
        partners |= 1u << q;
    }
  }
  auto pair_turns = [&](int q) -> const uint32_t* { return pair_ctr(q) + 3; };
  // The ring this streaming wave folds from: its own pair's, or pair 2j's when it takes the
  // second half of pair 2j's rows.
  const uint4* fmasks = masks;
  const uint32_t* fpub = ctr;
  if constexpr (kHalfShare) {
    if (share && half == 1) {
      fmasks = lds + (size_t)(pi - 1) * split_pair_u4(G, p.kl, D) + (size_t)G * lds_slots * 64;
      fpub = pair_ctr(pi - 1);
    }
  }

  // The streaming loop, in two instances: kWhole for runs whose steps all lie
  // in tiles wholly inside the matrix (every run but one that reaches a ragged
  // last tile), where a batch's rows sit at fixed distances from one uniform
  // base; otherwise each row is checked against n_rows and a row past it is
  // read from the zero buffer. Every batch slot is refilled unconditionally,
  // past the run's end with the run's last step again (rows in bounds, never
  // folded), so every path through the loop issues the same loads and the
  // compiler's wait counts keep all NB batches in flight: with the refill
  // skipped at the end and the per-row check in the same loop they came out as
  // waits for all but one or two of the batch's loads.
  // kShared: the pass has partners (see above), and the rows are loaded with the default
  // cache policy, so that a partner's reads of the same rows hit L2; without partners they
  // are loaded non-temporally, which streams faster alone (39,321,600 x 2 KiB, 3 keys in
  // one group: 12.3 against 13.1 ms; 8 keys in two groups: 16.1 against 22.1 ms, 1.06x
  // against 1.55-1.66x the matrix from HBM).
  auto consume = [&](auto whole_tag, auto shared_tag) -> bool {
    constexpr bool kWhole = decltype(whole_tag)::value;
    constexpr bool kShared = decltype(shared_tag)::value;
    auto ld = [](const uint4* a) -> uint4 {
      if constexpr (kShared)
        return *a;
      else
        return load_nt(a);
    };
    auto load = [&](uint64_t s, uint32_t b, uint4 (&v)[U][S]) {
      const uint64_t gs = gs0 + s;
      const uint64_t tile_base = (p.tile_begin + (gs >> (kl - 1))) << (kl + lb);
      const uint32_t pair = (uint32_t)(gs & (n_pairs - 1));
      const uint32_t r0 = b * U;
      if constexpr (kWhole) {
        const char* base = reinterpret_cast<const char*>(
            p.rows + (tile_base + ((uint64_t)(r0 >> 1) << kl) + 2 * (uint64_t)pair) * C);
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
          for (int j = 0; j < S; j++)
            v[u][j] = ld(reinterpret_cast<const uint4*>(base + row_off[u][j]));
      } else {
#pragma unroll
        for (int u = 0; u < U; u++) {
          const uint32_t r = r0 + u;
          const uint64_t x =
              tile_base + ((uint64_t)(r >> 1) << kl) + 2 * (uint64_t)pair + (r & 1);
          const uint4* src = x < p.n_rows ? p.rows + x * C : p.zero_rows;
#pragma unroll
          for (int j = 0; j < S; j++) v[u][j] = ld(src + lane_off[j]);
        }
      }
    };
    // The first NB batches go out in batch order, as the loop refills them
    // (the scheduler would otherwise be free to issue batch 0 last, and the wait
    // for it at the top of every step would become a wait for every load).
    uint4 ring[NB][U][S];
#pragma unroll
    for (int k = 0; k < NB; k++) {
      load(0, (uint32_t)k, ring[k]);
      __builtin_amdgcn_sched_barrier(0);
    }
    uint32_t buf = 0;  // step s's masks are in buffer s mod D
    uint32_t turn = 0;
    for (uint64_t s = 0; s < n_steps; ++s) {
      // Step s's masks are in the ring once every DFS wave of the pair has published it.
      if (!wait_count(fpub, (uint32_t)(s + 1))) return false;
      if (DW == 2 && !wait_count(ctr + 2, (uint32_t)(s + 1))) return false;
      const uint4* mbuf = fmasks + (size_t)buf * G * 128;
#pragma unroll
      for (uint32_t b0 = 0; b0 < n_batches; b0 += NB) {
        if constexpr (kShared) {
          // Publish this ring turn, then wait while a partner is more than kSplitPace
          // turns behind.
          ++turn;
          __hip_atomic_store(turns, turn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
          for (int q = 0; q < NP; q++) {
            if (!((partners >> q) & 1u)) continue;
            const uint32_t* tq = pair_turns(q);
            uint32_t spin = 0;
            while ((int32_t)(__hip_atomic_load(tq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) +
                             kSplitPace - turn) < 0) {
              if (++spin == kSplitSpinLimit) return false;
              __builtin_amdgcn_s_sleep(1);
            }
          }
        }
#pragma unroll
        for (int k = 0; k < NB; k++) {
          const uint32_t b = b0 + (uint32_t)k;
#pragma unroll
          for (int u = 0; u < U; u++)
#pragma unroll
            for (int g = 0; g < G; g++) {
              const uint4 m = mbuf[g * 128 + b * U + u];
#pragma unroll
              for (int j = 0; j < S; j++) xor_masked3(acc[g][j], ring[k][u][j], m);
            }
          // Batch b + NB goes into the freed slot: later in this step, or the
          // start of the next one (batch indices stay congruent mod NB).
          uint32_t nb = b + NB;
          uint64_t ns = s;
          if (nb >= n_batches) {
            nb -= n_batches;
            ++ns;
          }
          load(ns < n_steps ? ns : n_steps - 1, nb, ring[k]);
          // Rows of 2 or 4 chunks per lane: keep the refill ahead of the next batch's
          // folds. Left free, the scheduler sinks these loads to just before their use and
          // waits for each one (16,817,159 x 4 KiB, 1 key: 23.9 against 9.9 ms; 2 KiB rows,
          // 1 / 2 / 3 keys: 40.7 / 20.2 / 11.0 against 10.5 / 10.6 / 10.6 ms, one process).
          // Rows of one chunk per lane keep the same waits either way, and 3 keys over
          // 4 chunks per lane, whose streaming wave holds 256 VGPRs, would spill with it.
          if constexpr (S > 1 && !(S == 4 && G == 3)) __builtin_amdgcn_sched_barrier(0);
        }
      }
      lds_store_release(folded, (uint32_t)(s + 1));  // after this step's mask reads
      buf = buf + 1 == D ? 0 : buf + 1;
    }
    return true;
  };
  if (n_steps) {
    const uint64_t last_tile = p.tile_begin + ((gs0 + n_steps - 1) >> (kl - 1));
    using yes = std::integral_constant<bool, true>;
    using no = std::integral_constant<bool, false>;
    const bool whole = ((last_tile + 1) << (kl + lb)) <= p.n_rows;
    bool ok;
    if constexpr (kPaced)
      ok = partners ? (whole ? consume(yes{}, yes{}) : consume(no{}, yes{}))
                    : (whole ? consume(yes{}, no{}) : consume(no{}, no{}));
    else
      ok = whole ? consume(yes{}, no{}) : consume(no{}, no{});
    if (!ok) {
      atomicOr(p.fault, 1u);
      return;
    }
  }

  if (!live) return;
#pragma unroll
  for (int g = 0; g < G; g++) {
    if (key0 + g >= p.n_keys) continue;
#pragma unroll
    for (int j = 0; j < S; j++) {
      const uint32_t c = col0 + lane + 64 * j;
      if (c < col_end) p.partial[((size_t)part * p.n_keys + key0 + g) * C + c] = acc[g][j];
    }
  }
}

// -------------------------------------------------------- key evaluation --
// ChaChaKey::eval (dpf.rs:103-142) at n indices: one lane per index, a full
// d-level path walk each (only the index's low d bits are read).
__global__ __launch_bounds__(256) void eval_points(const DevKey* __restrict__ key,
                                                   const uint64_t* __restrict__ idx, uint64_t n,
                                                   uint4* __restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    out[i] = point_eval(key, idx[i]);
}

// Leaves [start, start + n_lanes * 2^kl) of the key's tree in order
// (ChaChaKey::eval_subtree / full_eval, dpf.rs:147-246), start aligned to
// 2^kl. Lane j walks d - kl levels from the root to the root of leaves
// [start + j*2^kl, +2^kl), then expands them depth-first with a per-lane LDS
// stack (kl levels of right children, as in the scan) and writes each leaf's
// 16 bytes in place.
__global__ __launch_bounds__(256) void eval_range(const DevKey* __restrict__ key, uint64_t start,
                                                  uint64_t n_lanes, int kl,
                                                  uint4* __restrict__ out) {
  extern __shared__ uint4 lds[];
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int d = (int)key->domain_bits;
  uint4* st_seed = lds + (size_t)wib * kl * 80;
  uint32_t* st_t = reinterpret_cast<uint32_t*>(st_seed + kl * 64);
  const int leaf_level = d - kl;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x; j < n_lanes; j += stride) {
    const uint64_t x0 = start + (j << kl);
    Node node = key_root(key);
    for (int level = 0; level < leaf_level; level++)
      node = ggm_step(key, level, node, (uint32_t)(x0 >> (d - 1 - level)) & 1u);
    uint4* o = out + (j << kl);
    if (kl == 0) {
      o[0] = leaf_mask(key, node);
      continue;
    }
    const uint32_t n_pairs = 1u << (kl - 1);
    int lvl = 0;
    for (uint32_t pair = 0; pair < n_pairs; ++pair) {
      if (pair) {
        lvl = kl - 1 - __builtin_ctz(pair);
        const uint4 sv = st_seed[lvl * 64 + lane];
        node.s0 = sv.x; node.s1 = sv.y; node.s2 = sv.z; node.s3 = sv.w;
        node.t = st_t[lvl * 64 + lane];
      }
      Node L, R;
      for (; lvl < kl - 1; ++lvl) {
        expand_node(key, leaf_level + lvl, node, L, R);
        st_seed[(lvl + 1) * 64 + lane] = make_uint4(R.s0, R.s1, R.s2, R.s3);
        st_t[(lvl + 1) * 64 + lane] = R.t;
        node = L;
      }
      expand_node(key, leaf_level + kl - 1, node, L, R);
      o[2 * pair] = leaf_mask(key, L);
      o[2 * pair + 1] = leaf_mask(key, R);
    }
  }
}

// ------------------------------------------------------------ reductions --
// out[key][c] (^)= XOR over parts of partial[part][key][c]. One block per
// (key, chunk group of 4); 256 threads stride over parts.
__global__ __launch_bounds__(256) void reduce_partials(const uint4* __restrict__ partial,
                                                       uint32_t parts, uint32_t n_keys,
                                                       uint32_t chunks, uint4* __restrict__ out,
                                                       const uint32_t* __restrict__ out_idx,
                                                       int accumulate) {
  __shared__ uint4 red[256];
  const uint32_t key = blockIdx.y;
  const uint32_t c = blockIdx.x;
  uint4 a = make_uint4(0, 0, 0, 0);
  for (uint32_t q = threadIdx.x; q < parts; q += blockDim.x)
    a = xor4(a, partial[((size_t)q * n_keys + key) * chunks + c]);
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = xor4(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    uint4* o = out + (size_t)(out_idx ? out_idx[key] : key) * chunks + c;
    *o = accumulate ? xor4(*o, red[0]) : red[0];
  }
}

// ----------------------------------------------------------------- fills --
// Word j of the matrix = splitmix64 output j for `seed`; 2 words per lane.
__global__ __launch_bounds__(256) void fill_splitmix(uint4* __restrict__ out, uint64_t first_pair,
                                                     uint64_t n_pairs, uint64_t seed) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pairs;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t v[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const uint64_t j = 2 * (first_pair + i) + h;
      uint64_t z = seed + (j + 1) * 0x9E3779B97F4A7C15ull;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      v[h] = z ^ (z >> 31);
    }
    out[i] = make_uint4((uint32_t)v[0], (uint32_t)(v[0] >> 32), (uint32_t)v[1],
                        (uint32_t)(v[1] >> 32));
  }
}

// chunked_matrix.rs:83-93 byte LCG. Each lane jumps to its 16-byte block.
__device__ __forceinline__ uint64_t lcg_jump(uint64_t seed, uint64_t steps) {
  uint64_t mul = 1, add = 0, cm = 6364136223846793005ull, ca = 1;
  while (steps) {
    if (steps & 1) {
      mul *= cm;
      add = add * cm + ca;
    }
    ca = ca * cm + ca;
    cm *= cm;
    steps >>= 1;
  }
  return mul * seed + add;
}

__global__ __launch_bounds__(256) void fill_lcg(uint4* __restrict__ out, uint64_t n_blocks16,
                                                uint64_t seed) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_blocks16;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t st = lcg_jump(seed, 16 * i);
    uint32_t wv[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      uint32_t word = 0;
#pragma unroll
      for (int b = 0; b < 4; b++) {
        st = st * 6364136223846793005ull + 1;
        word |= (uint32_t)(st >> 56) << (8 * b);
      }
      wv[q] = word;
    }
    out[i] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
  }
}

__global__ __launch_bounds__(256) void zero_u4(uint4* __restrict__ out, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = make_uint4(0, 0, 0, 0);
}

__global__ __launch_bounds__(256) void copy_u4(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                               uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// rows[idx[e]] ^= diffs[e]; one lane per 16-byte chunk. Duplicate indices
// are applied in one pass each by the host (see apply_delta), so no atomics.
__global__ __launch_bounds__(256) void xor_rows(uint4* __restrict__ rows, const uint64_t* __restrict__ idx,
                                                const uint4* __restrict__ diffs, uint64_t n_entries,
                                                uint32_t chunks) {
  const uint64_t total = n_entries * chunks;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = i / chunks;
    const uint32_t c = (uint32_t)(i % chunks);
    uint4* dst = rows + idx[e] * chunks + c;
    *dst = xor4(*dst, diffs[e * chunks + c]);
  }
}

__global__ __launch_bounds__(256) void scatter_u4(uint4* __restrict__ dst,
                                                  const uint4* __restrict__ src,
                                                  const uint32_t* __restrict__ idx, uint32_t n,
                                                  uint32_t per) {
  const uint64_t total = (uint64_t)n * per;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = i / per;
    const uint32_t j = (uint32_t)(i % per);
    dst[(uint64_t)idx[e] * per + j] = src[i];
  }
}

// ------------------------------------------------------------- launchers --

static int grid_for(uint64_t items, int block) {
  uint64_t g = (items + block - 1) / block;
  if (g > 65536 * 4) g = 65536 * 4;
  if (g == 0) g = 1;
  return (int)g;
}

size_t lane_dfs_lds_bytes(int waves_per_block, uint32_t kl, uint32_t gslots) {
  return (size_t)waves_per_block * (kl - 1 - gslots) * 64 * sizeof(uint4);
}

hipError_t launch_lane_tile_roots(const LaneDfsParams& p, uint4* seeds, uint32_t* ts,
                                  hipStream_t s) {
  const uint64_t n_tiles = p.tile_end - p.tile_begin;
  if (!n_tiles || !p.n_keys) return hipSuccess;
  if (p.n_keys > 65535 || (n_tiles + 255) / 256 > 0x7fffffffu) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lane_tile_roots, dim3((uint32_t)((n_tiles + 255) / 256), p.n_keys),
                     dim3(256), 0, s, p, seeds, ts);
  return hipGetLastError();
}

hipError_t launch_lane_dfs(int chunks, const LaneDfsParams& p, int waves_per_block,
                           hipStream_t s) {
  const uint32_t total = p.parts * p.n_keys;
  const dim3 block(64 * waves_per_block);
  const dim3 grid(xcd_grid(total, waves_per_block));
  if (p.gslots > p.kl - 1 || (p.gslots && !p.gstack)) return hipErrorInvalidValue;
  if ((p.root_seed != nullptr) != (p.root_t != nullptr) || (p.root_seed && !p.gslots))
    return hipErrorInvalidValue;
  const bool table = p.root_seed != nullptr;
  const size_t lds = lane_dfs_lds_bytes(waves_per_block, p.kl, p.gslots);
  switch (chunks) {
#define MG_LAUNCH(C, L, G, T)                                                       \
  hipLaunchKernelGGL((scan_lane_dfs<C, L, G, T>), grid, block, lds, s, p)
#define MG_CASE(C)                                                                  \
  case C:                                                                           \
    if (p.late_loads && p.gslots && table)                                          \
      MG_LAUNCH(C, true, true, true);                                               \
    else if (p.late_loads && p.gslots)                                              \
      MG_LAUNCH(C, true, true, false);                                              \
    else if (p.late_loads)                                                          \
      MG_LAUNCH(C, true, false, false);                                             \
    else if (p.gslots && table)                                                     \
      MG_LAUNCH(C, false, true, true);                                              \
    else if (p.gslots)                                                              \
      MG_LAUNCH(C, false, true, false);                                             \
    else                                                                            \
      MG_LAUNCH(C, false, false, false);                                            \
    break;
    MG_CASE(1) MG_CASE(2) MG_CASE(3) MG_CASE(4) MG_CASE(5) MG_CASE(6) MG_CASE(7) MG_CASE(8)
#undef MG_CASE
#undef MG_LAUNCH
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_suffix(int chunks, const SuffixParams& p, hipStream_t s) {
  const dim3 grid((uint32_t)((p.n_tiles + 3) / 4), p.n_keys);
  switch (chunks) {
#define MG_CASE(C)                                                            \
  case C:                                                                     \
    hipLaunchKernelGGL(scan_suffix<C>, grid, dim3(256), 0, s, p);             \
    break;
    MG_CASE(1) MG_CASE(2) MG_CASE(3) MG_CASE(4) MG_CASE(5) MG_CASE(6) MG_CASE(7) MG_CASE(8)
#undef MG_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// (S, G) instances of scan_split, each with NP = 1, 2 and 4 pairs per block. 4 keys at S = 4
// are left out: the streaming wave needs 256 VGPRs (it spills in 4-pair blocks), and at 1 or 2
// pairs per block only two pairs then fit per CU, which streams 19,660,800 x 4 KiB rows with
// 8 keys in 33.0 ms against 26.6 ms as groups of 3, 3 and 2.
#define MG_SPLIT_SHAPES(NN)                                                   \
  MG_SG(1, 1, NN) MG_SG(1, 2, NN) MG_SG(1, 3, NN) MG_SG(1, 4, NN)             \
  MG_SG(2, 1, NN) MG_SG(2, 2, NN) MG_SG(2, 3, NN) MG_SG(2, 4, NN)             \
  MG_SG(4, 1, NN) MG_SG(4, 2, NN) MG_SG(4, 3, NN)
#define MG_SPLIT_ALL MG_SPLIT_SHAPES(1) MG_SPLIT_SHAPES(2) MG_SPLIT_SHAPES(4)
// Two DFS waves per pair: 1 KiB rows, 2 or 4 keys.
#define MG_SPLIT_DW2 MG_SG2(1, 2, 1) MG_SG2(1, 2, 2) MG_SG2(1, 2, 4) \
  MG_SG2(1, 4, 1) MG_SG2(1, 4, 2) MG_SG2(1, 4, 4)

hipError_t launch_split(int S, int G, int NP, const StreamParams& p, hipStream_t s) {
  if (p.col_splits != 1 &&
      !(p.col_splits == 2 && p.chunks > 64u * S && p.chunks <= 128u * S && S <= 2))
    return hipErrorInvalidValue;
  const uint32_t n_groups = (p.n_keys + G - 1) / G;
  const uint32_t total = p.parts * n_groups * p.col_splits;
  const uint32_t blocks = (total + NP - 1) / NP;
  const uint32_t DW = p.dfs_waves;
  const dim3 block(64 * (DW + 1) * NP);
  const dim3 grid((blocks + kXcds - 1) / kXcds * kXcds);
  if (p.mask_depth < 2 || !p.fault) return hipErrorInvalidValue;
  const size_t lds = split_lds_bytes(G, p.kl, NP, p.mask_depth);
#define MG_SG(SS, GG, NN)                                                     \
  if (DW == 1 && S == SS && G == GG && NP == NN) {                            \
    hipLaunchKernelGGL((scan_split<SS, GG, NN, 1>), grid, block, lds, s, p);  \
    return hipGetLastError();                                                 \
  }
#define MG_SG2(SS, GG, NN)                                                    \
  if (DW == 2 && S == SS && G == GG && NP == NN) {                            \
    hipLaunchKernelGGL((scan_split<SS, GG, NN, 2>), grid, block, lds, s, p);  \
    return hipGetLastError();                                                 \
  }
  MG_SPLIT_ALL
  MG_SPLIT_DW2
#undef MG_SG
#undef MG_SG2
  return hipErrorInvalidValue;
}

size_t split_lds_bytes(int G, uint32_t kl, int NP, uint32_t mask_depth) {
  return (size_t)NP * split_pair_u4(G, kl, mask_depth) * sizeof(uint4);
}

hipError_t split_blocks_per_cu(int S, int G, int NP, int DW, size_t lds, int* blocks) {
#define MG_SG(SS, GG, NN)                                                                   \
  if (DW == 1 && S == SS && G == GG && NP == NN)                                            \
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, scan_split<SS, GG, NN, 1>,  \
                                                        128 * NN, lds);
#define MG_SG2(SS, GG, NN)                                                                  \
  if (DW == 2 && S == SS && G == GG && NP == NN)                                            \
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, scan_split<SS, GG, NN, 2>,  \
                                                        192 * NN, lds);
  MG_SPLIT_ALL
  MG_SPLIT_DW2
#undef MG_SG
#undef MG_SG2
  return hipErrorInvalidValue;
}

size_t stream_lds_bytes(int G, uint32_t kl, int waves_per_block) {
  return (size_t)waves_per_block * ((size_t)G * (kl - 1) * 64 + (size_t)G * 128) * sizeof(uint4);
}

// (S, G, PK) instances of scan_stream: PK > 1 only with S = 1.
#define MG_STREAM_SHAPES                                                      \
  MG_SG(1, 1, 1) MG_SG(1, 2, 1) MG_SG(1, 3, 1) MG_SG(1, 4, 1)                 \
  MG_SG(1, 1, 2) MG_SG(1, 2, 2) MG_SG(1, 3, 2) MG_SG(1, 4, 2)                 \
  MG_SG(1, 1, 4) MG_SG(1, 2, 4) MG_SG(1, 3, 4) MG_SG(1, 4, 4)                 \
  MG_SG(2, 1, 1) MG_SG(2, 2, 1) MG_SG(2, 3, 1)                                \
  MG_SG(4, 1, 1) MG_SG(4, 2, 1)

hipError_t launch_stream(int S, int G, int PK, const StreamParams& p, int waves_per_block,
                         hipStream_t s) {
  const uint32_t n_groups = (p.n_keys + G - 1) / G;
  const uint32_t total = p.parts * n_groups;
  const dim3 block(64 * waves_per_block);
  const dim3 grid(xcd_grid(total, waves_per_block));
  const size_t lds = stream_lds_bytes(G, p.kl, waves_per_block);
#define MG_SG(SS, GG, KK)                                                     \
  if (S == SS && G == GG && PK == KK) {                                       \
    hipLaunchKernelGGL((scan_stream<SS, GG, KK>), grid, block, lds, s, p);    \
    return hipGetLastError();                                                 \
  }
  MG_STREAM_SHAPES
#undef MG_SG
  return hipErrorInvalidValue;
}

hipError_t lane_dfs_blocks_per_cu(int chunks, bool late_loads, bool global_stack,
                                  bool root_table, int block_threads, size_t lds, int* blocks) {
  if (root_table && !global_stack) return hipErrorInvalidValue;
  switch (chunks) {
#define MG_OCC(C, L, G, T)                                                                 \
  hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, scan_lane_dfs<C, L, G, T>,          \
                                               block_threads, lds)
#define MG_CASE(C)                                                                          \
  case C:                                                                                   \
    if (late_loads)                                                                         \
      return global_stack ? (root_table ? MG_OCC(C, true, true, true)                       \
                                        : MG_OCC(C, true, true, false))                     \
                          : MG_OCC(C, true, false, false);                                  \
    return global_stack ? (root_table ? MG_OCC(C, false, true, true)                        \
                                      : MG_OCC(C, false, true, false))                      \
                        : MG_OCC(C, false, false, false);
    MG_CASE(1) MG_CASE(2) MG_CASE(3) MG_CASE(4) MG_CASE(5) MG_CASE(6) MG_CASE(7) MG_CASE(8)
#undef MG_CASE
#undef MG_OCC
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t stream_blocks_per_cu(int S, int G, int PK, int block_threads, size_t lds,
                                int* blocks) {
#define MG_SG(SS, GG, KK)                                                                   \
  if (S == SS && G == GG && PK == KK)                                                       \
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, scan_stream<SS, GG, KK>,    \
                                                        block_threads, lds);
  MG_STREAM_SHAPES
#undef MG_SG
  return hipErrorInvalidValue;
}

hipError_t launch_point(const PointParams& p, hipStream_t s) {
  const uint32_t entry_waves = (uint32_t)((p.n_entries + 63) / 64);
  const dim3 block(256);
  const dim3 grid((entry_waves + 3) / 4, p.n_keys);
  hipLaunchKernelGGL(scan_point, grid, block, 0, s, p);
  return hipGetLastError();
}

hipError_t launch_reduce(const uint4* partial, uint32_t parts, uint32_t n_keys, uint32_t chunks,
                         uint4* out, const uint32_t* out_idx, bool accumulate, hipStream_t s) {
  hipLaunchKernelGGL(reduce_partials, dim3(chunks, n_keys), dim3(256), 0, s, partial, parts,
                     n_keys, chunks, out, out_idx, accumulate ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_fill_splitmix(uint4* out, uint64_t first_pair, uint64_t n_pairs, uint64_t seed,
                                hipStream_t s) {
  hipLaunchKernelGGL(fill_splitmix, dim3(grid_for(n_pairs, 256)), dim3(256), 0, s, out,
                     first_pair, n_pairs, seed);
  return hipGetLastError();
}

hipError_t launch_fill_lcg(uint4* out, uint64_t n_blocks16, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(fill_lcg, dim3(grid_for(n_blocks16, 256)), dim3(256), 0, s, out, n_blocks16,
                     seed);
  return hipGetLastError();
}

hipError_t launch_copy(uint4* dst, const uint4* src, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(copy_u4, dim3(grid_for(n, 256)), dim3(256), 0, s, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_zero(uint4* out, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(zero_u4, dim3(grid_for(n, 256)), dim3(256), 0, s, out, n);
  return hipGetLastError();
}

hipError_t launch_scatter_u4(uint4* dst, const uint4* src, const uint32_t* idx, uint32_t n,
                             uint32_t per, hipStream_t s) {
  hipLaunchKernelGGL(scatter_u4, dim3(grid_for((uint64_t)n * per, 256)), dim3(256), 0, s, dst, src,
                     idx, n, per);
  return hipGetLastError();
}

hipError_t launch_xor_rows(uint4* rows, const uint64_t* idx, const uint4* diffs,
                           uint64_t n_entries, uint32_t chunks, hipStream_t s) {
  hipLaunchKernelGGL(xor_rows, dim3(grid_for(n_entries * chunks, 256)), dim3(256), 0, s, rows,
                     idx, diffs, n_entries, chunks);
  return hipGetLastError();
}

hipError_t launch_eval_points(const DevKey* key, const uint64_t* idx, uint64_t n, uint4* out,
                              hipStream_t s) {
  hipLaunchKernelGGL(eval_points, dim3(grid_for(n, 256)), dim3(256), 0, s, key, idx, n, out);
  return hipGetLastError();
}

hipError_t launch_eval_range(const DevKey* key, uint64_t start, uint64_t n_lanes, int kl,
                             uint4* out, hipStream_t s) {
  if (kl < 0 || kl > 12) return hipErrorInvalidValue;
  const size_t lds = (size_t)4 * kl * 80 * 16;
  hipLaunchKernelGGL(eval_range, dim3(grid_for(n_lanes, 256)), dim3(256), lds, s, key, start,
                     n_lanes, kl, out);
  return hipGetLastError();
}

}  // namespace mg
