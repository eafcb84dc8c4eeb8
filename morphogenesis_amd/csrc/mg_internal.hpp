// Host-side internals shared by the C ABI runtime (mg_runtime.cpp) and the
// request loop (mg_server.cpp): opaque handle layouts, error reporting,
// workspace buffers, launch planning and kernel enqueueing.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/morphogen_hip.h"
#include "mg_kernels.hpp"
#include "mg_key.hpp"

namespace mg {

// ---------------------------------------------------------------- errors --
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
const char* last_error();

#define MG_HIP(call)                                       \
  do {                                                     \
    hipError_t _e = (call);                                \
    if (_e != hipSuccess) return ::mg::hip_fail(_e, #call); \
  } while (0)

double now_ms();

// ------------------------------------------------------------ workspaces --
// Device buffer that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int reserve(size_t bytes);
  void release();
};

// Pinned host buffer that only grows.
struct HostBuf {
  void* p = nullptr;
  size_t cap = 0;
  int reserve(size_t bytes);
  void release();
};

// Launch plan of one scan; also the graph-cache signature (compared
// bytewise, so the constructor clears padding too).
struct Plan {
  uint32_t kernel;
  uint32_t chunks;
  uint32_t n_keys;
  uint32_t d, kl, lb, tb;
  uint32_t keys_per_wave;
  uint32_t waves_per_block;
  uint32_t parts;
  uint64_t tile_begin, tile_end, tiles_per_wave;
  uint64_t steps_per_wave;  // split kernel: leaf-pair steps per DFS/stream pair
  uint32_t gslots;          // lane-DFS: stack levels kept in global memory
  uint32_t late_loads;      // lane-DFS: rows loaded after a step's expansions
  uint32_t col_splits;      // split kernel: 1, or 2 pairs per row, one per 128-chunk half
  uint32_t mask_depth;      // split kernel: steps of masks a DFS wave may run ahead
  uint32_t dfs_waves;       // split kernel: DFS waves per pair (1 or 2)
  uint64_t row_begin, row_end;
  uint64_t n_delta;
  const void* rows;
  const void* delta_idx;
  const void* delta_rows;
  uint32_t delta_parts;
  // Buffers baked into a captured graph.
  const void* ws_keys;
  const void* ws_out;
  const void* ws_partial;
  const void* ws_partial_delta;
  const void* ws_gstack;
  const void* ws_key_idx;  // request loop: the launch's slot list
  Plan();
  bool operator==(const Plan& o) const;
};

// Device buffers one enqueue_scan reads and writes.
struct ScanBuffers {
  const DevKey* keys = nullptr;
  const uint32_t* key_idx = nullptr;  // launch key i -> keys[key_idx[i]]
  uint4* partial = nullptr;           // >= plan.parts * n_keys * chunks
  uint4* partial_delta = nullptr;     // >= plan.delta_parts * n_keys * chunks
  uint4* out = nullptr;
  const uint32_t* out_idx = nullptr;  // launch key i -> out row out_idx[i]
  bool accumulate = false;            // out ^= scan instead of out = scan
  uint4* gstack = nullptr;            // >= lane_gstack_bytes(plan, n_keys)
};

// One scan workspace: a stream of its own with the device buffers, pinned
// staging, events and graph cache that one call at a time uses on it
// (the role of a slot of the reference's ticket-based workspace pool,
// crates/morphogen-gpu-dpf/src/kernel.rs:50-72, 733-739). Calls on distinct
// workspaces run concurrently on their own streams.
struct Workspace {
  int index = 0;
  std::mutex mu;  // held by the call using this workspace
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {};
  DevBuf keys, partial, partial_delta, out;
  DevBuf gstack;  // lane-DFS stack levels kept in global memory
  DevBuf fault;   // the split kernel's hand-over timeout word for launches on `stream`
  HostBuf h_keys, h_out;
  // HIP graph of the kernel sequence for the last plan run with use_graph.
  bool graph_valid = false;
  Plan graph_plan;
  hipGraphExec_t graph_exec = nullptr;
  // Creates the stream, events and fault word on first use (device current).
  int ready();
  void invalidate_graph();
  void release();  // frees everything; the stream is synchronised first
};

// Readers-writer lock that lets a waiting writer in: a writer holds `turn_` while it waits
// for the readers in flight to leave, so new readers queue behind it instead of keeping it
// out for good. Works with std::shared_lock (readers) and std::unique_lock (writers).
class ScanGate {
 public:
  void lock() {
    std::lock_guard<std::mutex> t(turn_);
    rw_.lock();
  }
  void unlock() { rw_.unlock(); }
  void lock_shared() {
    { std::lock_guard<std::mutex> t(turn_); }
    rw_.lock_shared();
  }
  void unlock_shared() { rw_.unlock_shared(); }

 private:
  std::mutex turn_;
  std::shared_mutex rw_;
};

// Scan workspaces a device holds at most, and by default.
constexpr uint32_t kMaxWorkspaces = MG_MAX_WORKSPACES;
constexpr uint32_t kDefaultWorkspaces = 4;

}  // namespace mg

struct mg_device {
  int ordinal = 0;
  int cus = 0;
  size_t max_lds = 65536;  // dynamic LDS bytes one block may use
  // Control stream: matrix and delta writes, request-loop steps and the no-alloc
  // scan run here, each under `mu`.
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {};
  std::mutex mu;
  // Scans hold this shared for their whole call; writers of matrices and deltas (and a
  // request loop merging into the matrix it serves) hold it exclusively, after `mu`.
  mg::ScanGate state;
  // Scan workspaces (kMaxWorkspaces, streams created on first use); calls pick one of
  // the first `n_workspaces` by ticket.
  mg::Workspace ws[mg::kMaxWorkspaces];
  std::atomic<uint32_t> n_workspaces{mg::kDefaultWorkspaces};
  std::atomic<uint32_t> ticket{0};
  mg::DevBuf zero_rows;  // 4 KiB of zeros, read in place of rows past the end
  mg::DevBuf fault;      // the control stream's split-kernel timeout word
  std::mutex occ_mu;     // guards `occupancy`
  std::map<std::pair<uint32_t, uint32_t>, int> occupancy;  // (variant, lds) -> blocks/CU
  // mg_scan_noalloc: this device's address of the library's fixed scratch
  // array (set at open) and host room for MG_NOALLOC_MAX_KEYS parsed keys.
  void* noalloc_scratch = nullptr;
  std::vector<mg::DevKey> noalloc_keys;
  // Matrices, deltas and request loops still open on this device: closing
  // the device first is refused.
  std::atomic<int> children{0};
};

struct mg_matrix {
  mg_device* dev = nullptr;
  uint4* rows = nullptr;
  uint64_t n_rows = 0;
  uint32_t row_bytes = 0;
  // Request loops serving these rows (or holding them as a queued
  // snapshot): destroying the matrix first is refused.
  std::atomic<int> users{0};
};

struct mg_delta {
  mg_device* dev = nullptr;
  uint32_t row_bytes = 0;
  uint64_t capacity = 0;
  uint64_t n = 0;
  uint64_t epoch = 0;
  uint64_t max_idx = 0;
  uint64_t* d_idx = nullptr;
  uint4* d_diffs = nullptr;
  // Duplicate-free copy for the merge: each row index once, diffs XOR-ed.
  uint64_t n_unique = 0;
  uint64_t* d_uidx = nullptr;
  uint4* d_udiffs = nullptr;
};

namespace mg {

int set_device(mg_device* dev);

// Parses n_keys concatenated wire keys into `out`; all must share one domain.
int parse_keys(const uint8_t* keys, const size_t* key_lens, uint32_t n_keys, DevKey* out,
               uint32_t* common_d);

// Plans a scan of rows [o->row_begin, o->row_end) of a matrix at `rows`.
int make_plan(mg_device* dev, const void* rows, uint64_t n_rows, uint32_t row_bytes,
              uint32_t n_keys, uint32_t d, const mg_scan_opts* o, Plan* pl);

// Tile roots a lane-DFS launch with global stack levels precomputes (0: none, each
// wave walks its tile prefixes): one per key and tile, when the table is small.
inline uint64_t lane_root_entries(const Plan& pl, uint32_t n_keys) {
  const uint64_t e = (uint64_t)n_keys * (pl.tile_end - pl.tile_begin);
  return pl.gslots && n_keys <= 65535 && e * 20 <= kLaneRootTableMax ? e : 0;
}

// uint4s of the lane-DFS global region before the tile-root table: per wave, the
// stack levels and the tile-path cache.
inline size_t lane_wave_region_u4(const Plan& pl, uint32_t n_keys) {
  return (size_t)pl.parts * n_keys * ((size_t)pl.gslots * 64 + kLanePathU4);
}

// Bytes of the lane-DFS global region a plan's launch needs (0 if none): the waves'
// stack levels and tile-path caches, then the tile roots (seeds, then control words).
inline size_t lane_gstack_bytes(const Plan& pl, uint32_t n_keys) {
  if (!pl.gslots) return 0;
  const uint64_t e = lane_root_entries(pl, n_keys);
  return lane_wave_region_u4(pl, n_keys) * 16 + e * 16 + (e * 4 + 15) / 16 * 16;
}

// Where a launch sequence goes: the stream, and the word a split-kernel launch on it
// sets when a hand-over wait runs out.
struct Exec {
  hipStream_t stream = nullptr;
  uint32_t* fault = nullptr;
};
inline Exec control_exec(mg_device* dev) { return Exec{dev->stream, (uint32_t*)dev->fault.p}; }
inline Exec ws_exec(Workspace* w) { return Exec{w->stream, (uint32_t*)w->fault.p}; }

// Enqueues scan kernel, optional delta kernel and reductions on x.stream.
// Events (may be null) bracket the fused scan kernel.
int enqueue_scan(mg_device* dev, const Exec& x, const Plan& pl, const ScanBuffers& b,
                 hipEvent_t ev_scan0, hipEvent_t ev_scan1);

// Sets the plan's delta fields from opts->delta after checking it against
// the matrix and device.
int attach_delta(mg_device* dev, const mg_matrix* m, const mg_scan_opts* opts, Plan* pl);

// Fills *t from ev[0..4] (H2D, scan kernel, reductions, D2H) and the plan.
int fill_timing(const hipEvent_t ev[5], const Plan& pl, double t0, mg_timing* t);

// After x.stream has been synchronised: MG_ERR_DEVICE (and the word cleared) if a
// split-kernel launch since the last check reported a hand-over wait that ran out; a
// no-op for other plans.
int check_fault(const Exec& x, const Plan& pl);

// A workspace held for one call: the next free one of the device's first
// n_workspaces (by ticket), or, when all are busy, the ticket's one once free.
class WorkspaceLease {
 public:
  explicit WorkspaceLease(mg_device* dev);
  // Workspace `i`, waiting for it if busy.
  WorkspaceLease(mg_device* dev, uint32_t i);
  Workspace* operator->() const { return w_; }
  Workspace* get() const { return w_; }

 private:
  Workspace* w_ = nullptr;
  std::unique_lock<std::mutex> lk_;
};

// Drops every workspace's graph (they bake in matrix and delta addresses). Caller holds
// dev->state exclusively.
void invalidate_graphs(mg_device* dev);

// Resolves this device's address of the no-alloc scratch array (mg_noalloc.hip);
// called by mg_device_open with the device current.
int noalloc_setup(mg_device* dev);

// Collapses repeated row indices by XOR-ing their diffs (for the merge kernel).
void dedupe_delta(const uint64_t* idx, const uint8_t* diffs, uint64_t n, uint32_t row_bytes,
                  std::vector<uint64_t>* uidx, std::vector<uint8_t>* udiffs);

// Row width of the matrix a request loop serves (its answers' length).
uint32_t server_row_bytes(const mg_server* s);

}  // namespace mg
