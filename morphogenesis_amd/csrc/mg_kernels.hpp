// Kernel parameter blocks and launchers shared by the kernels and the host
// runtime. Host code sees only these declarations.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mg {

struct DevKey;

// Tree levels of a d-bit key are split as d = tb + lb + kl:
//   tb  wave-uniform tile prefix (cached across consecutive tiles),
//   lb  per-lane walk to the lane's subtree root (lanes >= 2^lb idle),
//   kl  depth of each lane's subtree, expanded depth-first (kl >= 1).
// A tile is 2^(lb + kl) consecutive rows.
// uint4s per wave of a lane-DFS launch's global region beyond its stack levels.
constexpr uint32_t kLanePathU4 = 64 + 16;

struct LaneDfsParams {
  const uint4* rows;
  uint64_t n_rows;
  const uint4* zero_rows;   // >= 16 zero chunks: read instead of rows past n_rows
  const DevKey* keys;
  const uint32_t* key_idx;  // key i of the launch is keys[key_idx[i]] (identity if null)
  uint32_t n_keys;
  uint32_t d, kl, lb, tb;
  uint64_t tile_begin;  // tiles [tile_begin, tile_end) are scanned; rows
  uint64_t tile_end;    // at or past n_rows are skipped
  // Wave w of part `part` takes leaf-pair steps [part * steps_per_wave, +steps_per_wave) of
  // the range, step = (tile - tile_begin) * 2^(kl-1) + pair; a run may start and end inside
  // a tile.
  uint64_t steps_per_wave;
  uint32_t parts;  // waves per key; wave w serves key w % n_keys
  uint4* partial;  // [parts][n_keys][chunks]
  // The top gslots levels of each lane's DFS stack (depths 1 .. gslots) live in
  // global memory, the rest in LDS. Wave w's region of gstack is gslots * 64 +
  // kLanePathU4 uint4s: [gslots][64] seeds, then the wave-uniform tile-prefix
  // path ([64] seeds, [64] control words) cached at tile starts.
  uint4* gstack;
  uint32_t gslots;
  uint32_t late_loads;  // rows loaded after a step's expansions (else before)
  // Global-stack instances: each key's tile roots (the node at level tb of every tile
  // of the range), [n_keys][tile_end - tile_begin] seeds and control words, from
  // launch_lane_tile_roots; null when each wave walks its tile prefixes itself.
  const uint4* root_seed;
  const uint32_t* root_t;
};

// Largest tile-root table a lane-DFS launch computes (bytes); a larger one is
// left out and each wave walks its tile prefixes itself.
constexpr size_t kLaneRootTableMax = 16u << 20;

struct StreamParams {
  const uint4* rows;
  uint64_t n_rows;
  uint32_t chunks;
  const DevKey* keys;
  const uint32_t* key_idx;
  uint32_t n_keys;
  uint32_t d, kl, lb, tb;
  uint64_t tile_begin;
  uint64_t tile_end;
  uint64_t tiles_per_wave;
  // scan_split: pair `part` takes leaf-pair steps [part * steps_per_wave, +steps_per_wave)
  // of the range, step = (tile - tile_begin) * 2^(kl-1) + pair; runs may start and end
  // inside a tile. scan_stream hands out whole tiles (tiles_per_wave).
  uint64_t steps_per_wave;
  uint32_t parts;  // waves per key group
  uint4* partial;  // [parts][n_keys][chunks]
  const uint4* zero_rows;  // >= 256 zero chunks, read in place of absent rows
  // scan_split: 1, or 2 when each pair takes one 64 x S-chunk column half of every row (rows
  // of 65-128 chunks with S = 1, 129-256 with S = 2); the halves write disjoint chunks of the
  // same partial slot.
  uint32_t col_splits;
  // scan_split: leaf-pair steps of masks a DFS wave may run ahead of its streaming wave
  // (the mask ring's depth, >= 2), and a device word the waves set when a hand-over wait
  // runs out (a bug guard: every wait is bounded, so the kernel always ends).
  uint32_t mask_depth;
  uint32_t* fault;
  // scan_split: DFS waves per pair (1, or 2 with the pair's keys split between them).
  uint32_t dfs_waves;
};

struct PointParams {
  const uint4* rows;        // matrix rows, or delta diffs when row_idx != 0
  const uint64_t* row_idx;  // delta row indices (nullptr for a matrix scan)
  uint64_t first_row;       // matrix scan: entry e is row first_row + e
  uint64_t n_entries;
  uint32_t chunks;
  const DevKey* keys;
  const uint32_t* key_idx;
  uint32_t n_keys;
  uint4* partial;  // [ceil(n_entries/64)][n_keys][chunks]
};

// Reference-shaped baseline (comparison only): 2^s-row tiles (s = min(11, d)),
// one prefix walk per tile, then an s-level path walk per row, as the
// reference's fused kernel does (fused_kernel.cu:309-360).
struct SuffixParams {
  const uint4* rows;
  uint64_t n_rows;
  const DevKey* keys;
  const uint32_t* key_idx;
  uint32_t n_keys;
  uint32_t d, s;
  uint64_t tile_begin, n_tiles;
  uint4* partial;  // [n_tiles][n_keys][chunks]
};

hipError_t launch_suffix(int chunks, const SuffixParams& p, hipStream_t s);
hipError_t launch_lane_dfs(int chunks, const LaneDfsParams& p, int waves_per_block,
                           hipStream_t s);
// Dynamic LDS of a lane-DFS block: kl - 1 - gslots levels of 64 seeds per wave.
size_t lane_dfs_lds_bytes(int waves_per_block, uint32_t kl, uint32_t gslots);
// Writes the tile roots of p's keys and tile range into seeds / ts ([n_keys][tiles]).
hipError_t launch_lane_tile_roots(const LaneDfsParams& p, uint4* seeds, uint32_t* ts,
                                  hipStream_t s);
hipError_t launch_stream(int S, int G, int PK, const StreamParams& p, int waves_per_block,
                         hipStream_t s);
size_t stream_lds_bytes(int G, uint32_t kl, int waves_per_block);
// The producer/consumer variant of the stream kernel: blocks of NP (1, 2, 4)
// pairs of a DFS wave and a row-streaming wave; needs lb = 6.
hipError_t launch_split(int S, int G, int NP, const StreamParams& p, hipStream_t s);
size_t split_lds_bytes(int G, uint32_t kl, int NP, uint32_t mask_depth);
hipError_t split_blocks_per_cu(int S, int G, int NP, int DW, size_t lds, int* blocks);
// Resident blocks per CU for the given block size and dynamic LDS.
hipError_t lane_dfs_blocks_per_cu(int chunks, bool late_loads, bool global_stack,
                                  bool root_table, int block_threads, size_t lds, int* blocks);
hipError_t stream_blocks_per_cu(int S, int G, int PK, int block_threads, size_t lds,
                                int* blocks);
hipError_t launch_point(const PointParams& p, hipStream_t s);
// out[out_idx ? out_idx[k] : k] (^)= XOR of the parts' partials of key k.
hipError_t launch_reduce(const uint4* partial, uint32_t parts, uint32_t n_keys, uint32_t chunks,
                         uint4* out, const uint32_t* out_idx, bool accumulate, hipStream_t s);
hipError_t launch_fill_splitmix(uint4* out, uint64_t first_pair, uint64_t n_pairs, uint64_t seed,
                                hipStream_t s);
hipError_t launch_fill_lcg(uint4* out, uint64_t n_blocks16, uint64_t seed, hipStream_t s);
hipError_t launch_zero(uint4* out, uint64_t n, hipStream_t s);
hipError_t launch_copy(uint4* dst, const uint4* src, uint64_t n, hipStream_t s);
// dst[idx[e] * per + j] = src[e * per + j] for e < n, j < per (16-byte units).
hipError_t launch_scatter_u4(uint4* dst, const uint4* src, const uint32_t* idx, uint32_t n,
                             uint32_t per, hipStream_t s);
hipError_t launch_xor_rows(uint4* rows, const uint64_t* idx, const uint4* diffs,
                           uint64_t n_entries, uint32_t chunks, hipStream_t s);
// out[i] = leaf of `key` at idx[i] (ChaChaKey::eval), i < n.
hipError_t launch_eval_points(const DevKey* key, const uint64_t* idx, uint64_t n, uint4* out,
                              hipStream_t s);
// out[i] = leaf start + i, i < n_lanes * 2^kl; start aligned to 2^kl, kl <= 12
// and kl <= d (ChaChaKey::eval_subtree).
hipError_t launch_eval_range(const DevKey* key, uint64_t start, uint64_t n_lanes, int kl,
                             uint4* out, hipStream_t s);

}  // namespace mg
