// Host runtime behind include/morphogen_hip.h: device, matrix, key, delta and
// scan entry points. It plays the role of the reference's GpuScanner /
// GpuPageMatrix host side (crates/morphogen-gpu-dpf/src/kernel.rs:156-1211,
// src/storage.rs:18-145) on one gfx950 GPU: a pool of scan workspaces, each a
// non-blocking HIP stream with grow-only device buffers, pinned staging,
// HIP-event timing and an optional HIP-graph replay of the launch sequence,
// so calls from several threads run on distinct streams; a control stream for
// matrix and delta writes.
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "mg_internal.hpp"

namespace mg {

namespace {
thread_local std::string g_err;
}

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? MG_ERR_OUT_OF_MEMORY : MG_ERR_DEVICE, "%s: %s (%d)",
              what, hipGetErrorString(e), (int)e);
}

const char* last_error() { return g_err.c_str(); }

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int DevBuf::reserve(size_t bytes) {
  if (bytes <= cap) return MG_OK;
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  const size_t want = std::max(bytes, (size_t)4096);
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    return hip_fail(e, "hipMalloc(workspace)");
  }
  cap = want;
  return MG_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

int HostBuf::reserve(size_t bytes) {
  if (bytes <= cap) return MG_OK;
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
  const size_t want = std::max(bytes, (size_t)4096);
  hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    p = nullptr;
    return hip_fail(e, "hipHostMalloc(staging)");
  }
  cap = want;
  return MG_OK;
}

void HostBuf::release() {
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
}

Plan::Plan() {
  memset(this, 0, sizeof *this);
  keys_per_wave = 1;
  waves_per_block = 4;
}

bool Plan::operator==(const Plan& o) const { return memcmp(this, &o, sizeof(Plan)) == 0; }

int Workspace::ready() {
  if (stream) return MG_OK;
  hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    stream = nullptr;
    return hip_fail(e, "hipStreamCreateWithFlags(workspace)");
  }
  int rc = MG_OK;
  for (auto& v : ev) {
    if ((e = hipEventCreate(&v)) != hipSuccess) {
      v = nullptr;
      rc = hip_fail(e, "hipEventCreate(workspace)");
      break;
    }
  }
  if (!rc) rc = fault.reserve(16);
  if (!rc && (launch_zero((uint4*)fault.p, 1, stream) != hipSuccess ||
              hipStreamSynchronize(stream) != hipSuccess))
    rc = fail(MG_ERR_DEVICE, "zeroing a workspace's fault word failed");
  if (rc) release();
  return rc;
}

void Workspace::invalidate_graph() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  graph_exec = nullptr;
  graph_valid = false;
}

void Workspace::release() {
  if (stream) (void)hipStreamSynchronize(stream);
  invalidate_graph();
  keys.release();
  partial.release();
  partial_delta.release();
  out.release();
  gstack.release();
  fault.release();
  h_keys.release();
  h_out.release();
  for (auto& v : ev) {
    if (v) (void)hipEventDestroy(v);
    v = nullptr;
  }
  if (stream) (void)hipStreamDestroy(stream);
  stream = nullptr;
}

WorkspaceLease::WorkspaceLease(mg_device* dev) {
  const uint32_t n = std::max(1u, std::min(dev->n_workspaces.load(), kMaxWorkspaces));
  const uint32_t t = dev->ticket.fetch_add(1) % n;
  for (uint32_t j = 0; j < n; ++j) {
    Workspace* w = &dev->ws[(t + j) % n];
    std::unique_lock<std::mutex> lk(w->mu, std::try_to_lock);
    if (lk.owns_lock()) {
      w_ = w;
      lk_ = std::move(lk);
      return;
    }
  }
  w_ = &dev->ws[t];
  lk_ = std::unique_lock<std::mutex>(w_->mu);
}

WorkspaceLease::WorkspaceLease(mg_device* dev, uint32_t i)
    : w_(&dev->ws[i]), lk_(dev->ws[i].mu) {}

void invalidate_graphs(mg_device* dev) {
  for (Workspace& w : dev->ws) {
    std::lock_guard<std::mutex> lk(w.mu);
    w.invalidate_graph();
  }
}

int set_device(mg_device* dev) {
  hipError_t e = hipSetDevice(dev->ordinal);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return MG_OK;
}

int parse_keys(const uint8_t* keys, const size_t* key_lens, uint32_t n_keys, DevKey* out,
               uint32_t* common_d) {
  size_t off = 0;
  for (uint32_t i = 0; i < n_keys; i++) {
    uint32_t party = 0;
    switch (parse_wire_key(keys + off, key_lens[i], &out[i], &party)) {
      case kParseOk:
        break;
      case kParseFormat:
        return fail(MG_ERR_KEY_FORMAT, "key %u: invalid key format (length %zu)", i, key_lens[i]);
      default:
        return fail(MG_ERR_DOMAIN, "key %u: domain bits %u > %d", i, (unsigned)keys[off + 18],
                    kMaxDomainBits);
    }
    if (i == 0) *common_d = out[i].domain_bits;
    if (out[i].domain_bits != *common_d)
      return fail(MG_ERR_DOMAIN, "key %u has domain %u, key 0 has %u; group keys by domain", i,
                  out[i].domain_bits, *common_d);
    off += key_lens[i];
  }
  return MG_OK;
}

void dedupe_delta(const uint64_t* idx, const uint8_t* diffs, uint64_t n, uint32_t row_bytes,
                  std::vector<uint64_t>* uidx, std::vector<uint8_t>* udiffs) {
  std::map<uint64_t, std::vector<uint8_t>> uniq;
  for (uint64_t i = 0; i < n; i++) {
    auto& v = uniq[idx[i]];
    if (v.empty()) v.assign(row_bytes, 0);
    const uint8_t* src = diffs + i * row_bytes;
    for (uint32_t b = 0; b < row_bytes; b++) v[b] ^= src[b];
  }
  uidx->clear();
  udiffs->clear();
  uidx->reserve(uniq.size());
  udiffs->reserve(uniq.size() * row_bytes);
  for (auto& kv : uniq) {
    uidx->push_back(kv.first);
    udiffs->insert(udiffs->end(), kv.second.begin(), kv.second.end());
  }
}

namespace {

// Resident waves the device holds for a kernel variant (cached per variant).
// Lane-DFS kinds pass G = late loads (1) + global stack levels (2) + tile-root
// table (4): the six kernels of one row width. Split kinds pass PK = DFS waves per pair and count
// two waves per pair, whatever the block's real size.
int resident_waves(mg_device* dev, uint32_t kind, int S, int G, int PK, int chunks,
                   int waves_per_block, size_t lds, int* out) {
  auto key = std::make_pair(((kind * 8 + (uint32_t)PK) * 8 + (uint32_t)G) * 64 +
                                (uint32_t)waves_per_block,
                            (uint32_t)lds);
  std::lock_guard<std::mutex> olk(dev->occ_mu);
  auto it = dev->occupancy.find(key);
  int blocks = 0;
  if (it != dev->occupancy.end()) {
    blocks = it->second;
  } else {
    if (kind == MG_KERNEL_LANE_DFS * 100 + (uint32_t)chunks)
      MG_HIP(lane_dfs_blocks_per_cu(chunks, (G & 1) != 0, (G & 2) != 0, (G & 4) != 0,
                                    64 * waves_per_block, lds, &blocks));
    else if (kind / 100 == MG_KERNEL_STREAM_SPLIT)
      MG_HIP(split_blocks_per_cu(S, G, waves_per_block / 2, PK, lds, &blocks));
    else
      MG_HIP(stream_blocks_per_cu(S, G, PK, 64 * waves_per_block, lds, &blocks));
    dev->occupancy[key] = blocks;
  }
  *out = std::max(1, blocks) * waves_per_block * dev->cus;
  return MG_OK;
}

// Keys per wave (stream) or per DFS/streaming pair (split) that a kernel instantiation takes
// at S chunks of 64 lanes per row: the split kernel's streaming wave holds 4 row batches and
// G x S accumulators (246 VGPRs for 4 KiB rows and 3 keys; 4 keys would spill).
int max_keys_per_wave(uint32_t kernel, int S) {
  if (kernel == MG_KERNEL_STREAM_SPLIT) return S <= 2 ? 4 : 3;
  return S == 1 ? 4 : (S == 2 ? 3 : 2);
}

// Split kernel: leaf-pair steps of masks a DFS wave may run ahead when the caller leaves it
// open (fewer when the block's LDS does not hold them; at least 2).
constexpr uint32_t kSplitDefaultDepth = 4;

// Split kernel: the most producer/consumer pairs per CU to keep resident, indexed by pairs
// per block (1, 2, 4). A fourth single-pair block fits the registers but does not pay at one
// key either: 78,643,200 x 1 KiB 11.62 against 11.28-11.32 ms, 19,660,800 x 4 KiB 12.14-12.17
// against 12.32 ms (same box).
constexpr int kSplitPairsPerCu[5] = {0, 3, 4, 0, 4};

int stream_s(uint32_t chunks) { return chunks <= 64 ? 1 : (chunks <= 128 ? 2 : 4); }

// The split kernel's chunks per lane group: a pair that takes one 128-chunk column half of
// each row works as on 2 KiB rows.
int split_s(const Plan* pl) {
  return pl->col_splits == 2 ? (pl->chunks > 128 ? 2 : 1) : stream_s(pl->chunks);
}

// Keys per pass from which rows of 129-256 chunks go in column halves by default. Over
// 19,660,800 x 4 KiB rows: 4 keys 11.55 against 16.44 ms (one group of 4 per half against
// 2 + 2), 8 keys 18.06 against 26.46 ms; 12 keys 29.8 ms both ways; 1-3 keys 11.27-11.30
// against 11.05-11.25 ms. 2 KiB rows keep whole rows: there a pair holds 4 keys either way,
// and the halves only double the DFS work (39,321,600 x 2 KiB, 4 / 8 / 12 keys: 13.68 / 23.17
// / 37.37 against 11.83 / 18.72 / 27.99 ms).
constexpr uint32_t kSplitHalvesFromKeys = 4;

// Rows of 129-256 chunks in column halves by default: two pairs per row, each holding up to
// 4 keys' accumulators of its half, in place of one pair with at most 3 keys.
bool split_halves_by_default(uint32_t chunks, uint32_t n_keys) {
  return chunks > 128 && n_keys >= kSplitHalvesFromKeys;
}

// Rows per load instruction of the stream kernel: rows of at most 16 (32)
// chunks go 4 (2) to a load. A 144-byte row then fills 36 lanes of a load
// instead of 9.
int stream_pk(uint32_t chunks) { return chunks <= 16 ? 4 : (chunks <= 32 ? 2 : 1); }

// Keys per wave (stream) or per block (split) when the caller leaves it open. Keys go evenly
// into the fewest groups of at most gmax keys: a pass costs about as much as its widest group,
// so 4 keys run as 2 + 2 (32 GiB of 272-byte rows, stream kernel: 44.0 ms as 3 + 1). The split
// kernel's cost per key falls up to its widest group (1 KiB rows: 11.7 ms for one key, 14.9 ms
// for four), the stream kernel's up to 3.
int default_keys_per_wave(uint32_t kernel, int S, uint32_t n_keys) {
  const int gmax = kernel == MG_KERNEL_STREAM_SPLIT ? max_keys_per_wave(kernel, S)
                                                    : std::min(3, max_keys_per_wave(kernel, S));
  const int nk = std::max(1, (int)n_keys);
  const int ng = (nk + gmax - 1) / gmax;
  // Split passes of several groups over rows of two chunks per lane (2 KiB rows, column
  // halves of 4 KiB rows) pace their pairs, and full groups run faster there than keys
  // spread evenly (39,321,600 x 2 KiB, 5-6 keys: 15.5 ms as 4 + 1-2 against 17.9 ms as
  // 3 + 2-3; 19,660,800 x 4 KiB: 16.1 against 18.2-19.6 ms).
  if (kernel == MG_KERNEL_STREAM_SPLIT && S == 2 && ng > 1) return gmax;
  return (nk + ng - 1) / ng;
}

// Splits a d-bit domain as tb + lb + kl (see LaneDfsParams).
void split_levels(uint32_t d, uint32_t kl_req, Plan* pl) {
  const uint32_t kl = std::min(std::max(kl_req, 1u), d);
  const uint32_t lb = std::min(6u, d - kl);
  pl->kl = kl;
  pl->lb = lb;
  pl->tb = d - kl - lb;
}

// Lane-DFS launch shape. Blocks of `wpb` waves are dealt to the CUs evenly and
// their waves to the 4 SIMDs, so the most loaded SIMD runs ceil(waves / SIMDs)
// waves. A wave takes an equal run of leaf-pair steps (2^(kl-1) per tile,
// runs may start and end inside a tile): per lane about two node expansions
// per step, 8 more per tile it touches (the lb-level path to the lane root and
// the tile prefix) and kl - 1 to seek its first step. The SIMD's rate depends
// on how many waves share it, relative to 8 waves: the node routine alone
// measured 0.488, 0.834, 0.878, 0.908, 0.946, 0.966, 0.986, 1.0 at 1 .. 8 waves
// (the stage-spaced s_nops leave a lone wave idle, more waves fill them); the
// whole kernel loses more at 3-4 waves (2^27 x 32 B, one key, kl = 9: 0.499,
// 0.827, 0.869 and 1.0 of the 7-wave rate at about 1, 2, 3.5 and 7 waves per
// SIMD), which the table below follows. The shape minimising the most loaded
// SIMD's work / rate wins; kl is searched when `kl_fixed` is 0, waves per key
// when `parts_fixed` is 0.
// Lane-DFS stack levels the plan may keep in global memory instead of LDS.
constexpr uint32_t kLaneMaxGlobalSlots = 4;

struct LaneShape {
  uint32_t kl = 0;
  uint32_t wpb = 0;
  uint32_t gslots = 0;
  uint32_t late_loads = 0;
  uint64_t parts = 0;
  double cost = 0;
};

int choose_lane_shape(mg_device* dev, const Plan& pl0, uint32_t d, uint32_t n_keys, uint32_t kl_fixed,
                      uint32_t wpb_req, uint64_t parts_fixed, uint32_t gs_fixed,
                      uint32_t loads_fixed, LaneShape* out) {
  static const double kRate[9] = {0, 0.49, 0.82, 0.84, 0.86, 0.90, 0.94, 0.986, 1.0};
  const uint64_t simds = (uint64_t)dev->cus * 4;
  LaneShape best;
  const uint32_t kl_lo = kl_fixed ? std::min(kl_fixed, d) : 1;
  const uint32_t kl_hi = kl_fixed ? kl_lo : std::min(10u, d);
  for (uint32_t kl = kl_lo; kl <= kl_hi; ++kl) {
    const uint32_t lb = std::min(6u, d - kl);
    const uint64_t tile_rows = 1ull << (kl + lb);
    // A range start off this depth's tiles rules it out; a fixed depth is kept, so that the
    // plan's own alignment check reports it.
    if (!kl_fixed && pl0.row_begin % tile_rows != 0) continue;
    const uint64_t n_tiles = std::max<uint64_t>(
        1, (pl0.row_end + tile_rows - 1) / tile_rows - pl0.row_begin / tile_rows);
    // Stack levels moved to global memory: none, or as many as lift the waves per CU.
    // (A fixed count beyond this depth's levels is reported by the plan.)
    const uint32_t gs_lo = gs_fixed ? std::min(gs_fixed - 1, kl - 1) : 0;
    const uint32_t gs_hi = gs_fixed ? gs_lo : std::min(kLaneMaxGlobalSlots, kl - 1);
    for (uint32_t gs = gs_lo; gs <= gs_hi; ++gs) {
      for (uint32_t late = 0; late <= 1; ++late) {
        if (loads_fixed && late != loads_fixed - 1) continue;
        uint32_t wpb = wpb_req;
        while (wpb > 1 && lane_dfs_lds_bytes((int)wpb, kl, gs) > dev->max_lds) wpb--;
        const size_t lds = lane_dfs_lds_bytes((int)wpb, kl, gs);
        if (lds > dev->max_lds) continue;
        // With global levels the launch precomputes the tile roots when the table is small.
        const bool table = gs && n_keys <= 65535 && n_keys * n_tiles * 20 <= kLaneRootTableMax;
        int slots = 0;
        int rc = resident_waves(dev, MG_KERNEL_LANE_DFS * 100 + pl0.chunks, 0,
                                (int)late + (gs ? 2 : 0) + (table ? 4 : 0), 1,
                                (int)pl0.chunks, (int)wpb, lds, &slots);
        if (rc) return rc;
        const uint64_t cap = std::max<uint64_t>((uint64_t)slots, 1);
        // Candidate waves per key: the most that stay resident and a few fewer, the counts that
        // fill a whole number of waves per SIMD, and those of whole tiles per wave.
        const uint64_t n_pairs = 1ull << (kl - 1);
        const uint64_t total_steps = n_tiles * n_pairs;
        uint64_t cand[32];
        int nc = 0;
        if (parts_fixed) {
          cand[nc++] = parts_fixed;
        } else {
          const uint64_t p_hi = std::max<uint64_t>(1, cap / n_keys);
          for (uint64_t j = 0; j < 9 && j < p_hi; ++j) cand[nc++] = p_hi - j;
          for (uint64_t m = 1; m <= 8; ++m)
            if (m * simds >= n_keys) cand[nc++] = m * simds / n_keys;
          const uint64_t t_lo = std::max<uint64_t>(1, (n_tiles * n_keys + cap - 1) / cap);
          for (uint64_t t = t_lo; t < t_lo + 8; ++t) cand[nc++] = (n_tiles + t - 1) / t;
        }
        for (int ci = 0; ci < nc; ++ci) {
          const uint64_t p0 = std::max<uint64_t>(1, std::min(cand[ci], total_steps));
          const uint64_t spw = (total_steps + p0 - 1) / p0;
          const uint64_t parts = (total_steps + spw - 1) / spw;
          const uint64_t waves = parts * n_keys;
          // Waves beyond the resident ones run in later rounds.
          const uint64_t rounds = (waves + cap - 1) / cap;
          const uint64_t per_round = (waves + rounds - 1) / rounds;
          const uint64_t wps = (per_round + simds - 1) / simds;
          const double rate = kRate[std::min<uint64_t>(wps, 8)];
          // Per lane: two expansions per step, 8 per tile touched (6 with the tile-root
          // table), the seek when a run starts inside a tile, and without the table the
          // wave's first walk down the tile prefix.
          const bool whole_tiles = spw % n_pairs == 0;
          const double tiles_touched = (double)((spw + n_pairs - 1) / n_pairs + (whole_tiles ? 0 : 1));
          const double work = 2.0 * (double)spw + (table ? 6.0 : 8.0) * tiles_touched +
                              (whole_tiles ? 0.0 : kl - 1.0) +
                              (table ? 0.0 : (double)(d - kl - lb));
          // Rows loaded after a step's expansions wait in the open: 1 % dearer at
          // equal waves per SIMD (2^27 x 32 B, one key, four waves per SIMD: 2.87
          // against 2.83 ms), so they win only where they add waves.
          const double cost =
              (double)rounds * (double)wps * work / rate * (late ? 1.01 : 1.0);
          if (best.kl == 0 || cost < best.cost * 0.999) {
            best.kl = kl;
            best.wpb = wpb;
            best.gslots = gs;
            best.late_loads = late;
            best.parts = parts;
            best.cost = cost;
          }
        }
      }
    }
  }
  if (best.kl == 0)
    return fail(MG_ERR_ROWS, "row_begin %llu is not a multiple of any lane-DFS tile",
                (unsigned long long)pl0.row_begin);
  *out = best;
  return MG_OK;
}

}  // namespace

int make_plan(mg_device* dev, const void* rows, uint64_t n_rows, uint32_t row_bytes,
              uint32_t n_keys, uint32_t d, const mg_scan_opts* o, Plan* pl) {
  mg_scan_opts def{};
  if (!o) o = &def;
  pl->chunks = row_bytes / 16;
  pl->n_keys = n_keys;
  pl->d = d;
  pl->rows = rows;
  pl->row_begin = o->row_begin;
  pl->row_end = o->row_end ? o->row_end : n_rows;
  if (pl->row_end > n_rows || pl->row_begin > pl->row_end)
    return fail(MG_ERR_ROWS, "row range [%llu, %llu) outside the %llu-row matrix",
                (unsigned long long)pl->row_begin, (unsigned long long)pl->row_end,
                (unsigned long long)n_rows);
  if (d < 64 && (n_rows > (1ull << d)))
    return fail(MG_ERR_ROWS, "%llu rows exceed the 2^%u-leaf key domain",
                (unsigned long long)n_rows, d);
  if (o->waves_per_block > 8 ||
      (o->waves_per_block > 4 && o->kernel != MG_KERNEL_STREAM_SPLIT))
    return fail(MG_ERR_INVALID_ARG, "waves_per_block must be 0-4 (split kernel: 2, 4 or 8)");
  if (o->keys_per_wave > 4) return fail(MG_ERR_INVALID_ARG, "keys_per_wave must be 0-4");
  if (o->lane_depth > 16) return fail(MG_ERR_INVALID_ARG, "lane_depth must be 0-16");
  if (o->mask_depth == 1 || o->mask_depth > 8)
    return fail(MG_ERR_INVALID_ARG, "mask_depth must be 0 (tuned) or 2-8");

  uint32_t kernel = o->kernel;
  if (kernel > MG_KERNEL_STREAM_SPLIT)
    return fail(MG_ERR_INVALID_ARG, "unknown kernel %u", kernel);
  if (kernel == MG_KERNEL_AUTO) {
    // Split against stream over 32 GiB (measured in one process): the split kernel wins every
    // key count at 1 KiB rows (1 / 3 / 8 keys: 5.08 / 5.79 / 11.77 ms against 5.32 / 6.33 /
    // 18.74 ms) and three or more keys at 528-1008-byte rows (768 B, 3 keys: 7.30 against
    // 8.24 ms), but loses one or two keys there (768 B, 1 key: 6.04 against 5.26 ms) and
    // everywhere below 528 B, where rows pack several to a load. Over 80.5 GB of 2 KiB and
    // 4 KiB rows it wins every key count, most where its groups hold more keys than a stream
    // wave (4 KiB, 1 / 3 / 8 keys: 11.74 / 11.79 / 26.2 ms against 12.56 / 23.4 / 46.4 ms;
    // 2 KiB, 1 / 8 keys: 12.75 / 19.0 ms against 13.64 / 38.0 ms). It needs d >= kl + 6; its
    // default kl (5 to 7) shrinks to d - 6 on smaller domains.
    const uint32_t kl_split = o->lane_depth ? o->lane_depth : 5;
    if (d == 0) kernel = MG_KERNEL_POINT;
    else if (pl->chunks <= 8) kernel = MG_KERNEL_LANE_DFS;
    else if (pl->chunks > 32 && pl->chunks <= 256 && (n_keys >= 3 || pl->chunks >= 64) &&
             d >= kl_split + 6)
      kernel = MG_KERNEL_STREAM_SPLIT;
    else if (pl->chunks <= 256) kernel = MG_KERNEL_STREAM;
    else kernel = MG_KERNEL_POINT;
  }
  if (kernel != MG_KERNEL_POINT && d == 0)
    return fail(MG_ERR_INVALID_ARG, "a 0-bit domain needs the point kernel");
  if (kernel == MG_KERNEL_LANE_DFS && pl->chunks > 8)
    return fail(MG_ERR_INVALID_ARG, "lane-DFS kernel takes rows of at most 128 bytes");
  if ((kernel == MG_KERNEL_STREAM || kernel == MG_KERNEL_STREAM_SPLIT) && pl->chunks > 256)
    return fail(MG_ERR_INVALID_ARG, "stream kernels take rows of at most 4096 bytes");
  if (kernel == MG_KERNEL_SUFFIX && (pl->chunks > 8 || d == 0))
    return fail(MG_ERR_INVALID_ARG, "suffix-walk kernel takes rows of at most 128 bytes, d >= 1");
  pl->kernel = kernel;
  pl->waves_per_block = o->waves_per_block ? o->waves_per_block : 4;
  pl->col_splits = 1;
  if (o->row_halves > 2)
    return fail(MG_ERR_INVALID_ARG, "row_halves %u: 0 (tuned), 1 (whole rows) or 2 (halves)",
                o->row_halves);
  if (kernel == MG_KERNEL_STREAM_SPLIT) {
    if (o->row_halves == 2 && pl->chunks <= 64)
      return fail(MG_ERR_INVALID_ARG, "column halves need rows of more than 1 KiB");
    pl->col_splits = (o->row_halves == 2 ||
                      (o->row_halves == 0 && split_halves_by_default(pl->chunks, n_keys)))
                         ? 2 : 1;
  }

  if (kernel == MG_KERNEL_POINT) {
    pl->parts = (uint32_t)((pl->row_end - pl->row_begin + 63) / 64);
    return MG_OK;
  }
  if (kernel == MG_KERNEL_SUFFIX) {
    pl->lb = std::min<uint32_t>(11, d);  // suffix levels
    const uint64_t tile_rows = 1ull << pl->lb;
    if (pl->row_begin % tile_rows != 0)
      return fail(MG_ERR_ROWS, "row_begin %llu is not a multiple of the %llu-row tile",
                  (unsigned long long)pl->row_begin, (unsigned long long)tile_rows);
    pl->tile_begin = pl->row_begin / tile_rows;
    pl->tile_end = (pl->row_end + tile_rows - 1) / tile_rows;
    const uint64_t n_tiles = pl->tile_end - pl->tile_begin;
    if (n_tiles >= (1ull << 32)) return fail(MG_ERR_ROWS, "too many tiles");
    pl->parts = (uint32_t)std::max<uint64_t>(n_tiles, 1);
    return MG_OK;
  }
  // Default lane depth (measured on MI355X). Lane-DFS: the cost model of choose_lane_shape
  // over kl and tiles per wave (2^31 x 32 B, 24 keys: kl = 9, 921 ms; 935 / 937 / 940 ms at
  // kl = 8 / 10 / 11. 2^27 x 32 B: 1 key 2.56 ms at kl = 9 against 3.25 ms at kl = 6; 2 keys
  // 4.94 ms at kl = 10; 3 keys 7.35 ms at kl = 7). The stream kernel at 78,643,200 x 1 KiB:
  // one key per wave runs best at kl = 4 (11.97 ms; 12.08 ms at kl = 6), two keys per wave at
  // kl = 6 (13.04 ms; 13.14 ms at kl = 7), three at kl = 4 to 7 alike (15.02 to 15.13 ms).
  // Narrower rows carry more PRG work per byte and want fewer lane-root walks: 32 GiB of
  // 144-byte rows, one key, 10.4 ms at kl = 4 and 9.2 ms at kl = 6; 272-byte rows 7.04 / 6.66
  // ms; 512-byte rows 5.36 / 5.25 ms.
  uint32_t kl_default;
  LaneShape lane_shape;
  if (kernel == MG_KERNEL_LANE_DFS) {
    if (o->stack_global > kLaneMaxGlobalSlots + 1)
      return fail(MG_ERR_INVALID_ARG, "stack_global %u: at most %u stack levels in global memory",
                  o->stack_global, kLaneMaxGlobalSlots);
    if (o->row_loads > 2)
      return fail(MG_ERR_INVALID_ARG, "row_loads %u: 0 (tuned), 1 (before) or 2 (after)",
                  o->row_loads);
    int rc0 = choose_lane_shape(dev, *pl, d, n_keys, o->lane_depth,
                                (uint32_t)pl->waves_per_block, o->waves_per_key, o->stack_global,
                                o->row_loads, &lane_shape);
    if (rc0) return rc0;
    kl_default = lane_shape.kl;
  } else if (kernel == MG_KERNEL_STREAM_SPLIT) {
    // Split kernel (kernel time, one box, the default mask ring): rows of up to 64 chunks
    // want deep lane subtrees, fewer as a group's keys add stack LDS: 78,643,200 x 1 KiB,
    // 1 key 11.07 / 11.15 / 11.27 ms at kl = 9 / 8 / 7; 2 keys 11.29 / 11.43 ms at kl = 8 / 7;
    // 3 keys 11.46 / 11.54 / 11.88 ms at kl = 8 / 9 / 6. Four-key groups hold five stack
    // levels in registers, which leaves room at kl = 9 for a 3-step mask ring: 12.88 /
    // 13.10 / 13.15 / 13.35 ms at kl = 9 / 8 / 7 / 6 (rings of 3 / 3 / 4 / 4 steps; kl = 9
    // with 2 steps 13.68 ms).
    // Wider rows stream best at kl = 7 (19,660,800 x 4 KiB, 1 key 12.28 against 12.38 ms at
    // kl = 8; 3 keys 12.27 against 12.30 / 12.37 ms at kl = 6 / 8; 39,321,600 x 2 KiB, 3 keys
    // 12.44 against 12.56 / 12.58 ms at kl = 6 / 8).
    const int S = split_s(pl);
    const int g = o->keys_per_wave ? (int)o->keys_per_wave : default_keys_per_wave(kernel, S, n_keys);
    kl_default = S > 1 ? 7 : (g == 3 ? 8 : 9);
    if (d > 6) kl_default = std::min(kl_default, d - 6);
  } else {
    const int g = o->keys_per_wave ? (int)o->keys_per_wave
                                   : default_keys_per_wave(kernel, stream_s(pl->chunks), n_keys);
    kl_default = g <= 2 ? 6 : 7;
  }
  split_levels(d, o->lane_depth ? o->lane_depth : kl_default, pl);
  if (kernel == MG_KERNEL_STREAM_SPLIT && pl->lb != 6)
    return fail(MG_ERR_INVALID_ARG,
                "the split stream kernel needs d >= kl + 6 (d = %u, kl = %u)", d, pl->kl);
  const uint64_t tile_rows = 1ull << (pl->kl + pl->lb);
  if (pl->row_begin % tile_rows != 0)
    return fail(MG_ERR_ROWS, "row_begin %llu is not a multiple of the %llu-row tile",
                (unsigned long long)pl->row_begin, (unsigned long long)tile_rows);
  pl->tile_begin = pl->row_begin / tile_rows;
  pl->tile_end = (pl->row_end + tile_rows - 1) / tile_rows;
  const uint64_t n_tiles = pl->tile_end - pl->tile_begin;

  uint32_t groups = n_keys;
  int resident = 0;
  int rc = MG_OK;
  if (kernel == MG_KERNEL_LANE_DFS) {
    if (!o->waves_per_block) pl->waves_per_block = lane_shape.wpb;
    if (o->stack_global) {
      if (o->stack_global - 1 > pl->kl - 1)
        return fail(MG_ERR_INVALID_ARG, "stack_global %u: a depth-%u lane stack has %u levels",
                    o->stack_global, pl->kl, pl->kl - 1);
      pl->gslots = o->stack_global - 1;
    } else {
      // The shape's global stack levels apply to its own depth only.
      pl->gslots = lane_shape.kl == pl->kl ? lane_shape.gslots : 0;
    }
    pl->late_loads = o->row_loads ? o->row_loads - 1
                                  : (lane_shape.kl == pl->kl ? lane_shape.late_loads : 0);
    const size_t lds = lane_dfs_lds_bytes((int)pl->waves_per_block, pl->kl, pl->gslots);
    if (lds > dev->max_lds)
      return fail(MG_ERR_INVALID_ARG, "lane-DFS plan needs %zu B of LDS per block (max %zu)", lds,
                  dev->max_lds);
    rc = resident_waves(dev, MG_KERNEL_LANE_DFS * 100 + pl->chunks, 0,
                        (int)pl->late_loads + (pl->gslots ? 2 : 0) +
                            (lane_root_entries(*pl, n_keys) ? 4 : 0),
                        1, (int)pl->chunks, (int)pl->waves_per_block, lds, &resident);
  } else {
    const int S = kernel == MG_KERNEL_STREAM_SPLIT ? split_s(pl) : stream_s(pl->chunks);
    int G = o->keys_per_wave ? (int)o->keys_per_wave : default_keys_per_wave(kernel, S, n_keys);
    G = std::max(1, std::min(G, max_keys_per_wave(kernel, S)));
    pl->keys_per_wave = (uint32_t)G;
    groups = (n_keys + G - 1) / G;
    if (kernel == MG_KERNEL_STREAM) {
      if (!o->waves_per_block) pl->waves_per_block = 1;
      const size_t lds = stream_lds_bytes(G, pl->kl, (int)pl->waves_per_block);
      if (lds > dev->max_lds)
        return fail(MG_ERR_INVALID_ARG, "stream plan needs %zu B of LDS per block (max %zu)",
                    lds, dev->max_lds);
      rc = resident_waves(dev, MG_KERNEL_STREAM * 100 + (uint32_t)(S * 8 + G), S, G,
                          stream_pk(pl->chunks), (int)pl->chunks, (int)pl->waves_per_block, lds,
                          &resident);
      // With two or more keys per wave, more than two waves per SIMD run slower (32 GiB of
      // 272-byte rows, 2 keys: 10.15 ms at 1928 waves, 13.5 ms at 2804); one key per wave
      // gains from the third (6.58 against 7.21 ms).
      if (G >= 2) resident = std::min(resident, 2 * 4 * dev->cus);
    }
  }
  if (kernel == MG_KERNEL_STREAM_SPLIT) {
    // A block holds NP pairs of a DFS wave and a streaming wave; parts count pairs.
    const int S = split_s(pl);
    const int G = (int)pl->keys_per_wave;
    // Default: four pairs per block from three keys per group on, when their LDS fits. A
    // block's eight waves then put one DFS wave on each SIMD; single-pair blocks, three to a
    // CU, can stack two DFS waves on one SIMD (78,643,200 x 1 KiB, 3 keys: 12.90 against
    // 13.06 ms; 4 keys: 14.65 against 15.15 ms). Two pairs per block measured slower than
    // either. One or two keys per group: single-pair blocks (1 key: 11.39 against 11.51 ms),
    // unless the pass has several groups: a 4-pair block then holds one part's pairs of four
    // groups, which read the same rows at about the same time and share them in L2
    // (19,660,800 x 4 KiB, 4 keys as 2 + 2: 16.53 against 20.80 ms).
    uint32_t wpb = o->waves_per_block;
    if (wpb != 2 && wpb != 4 && wpb != 8) {
      // 0, or a block size meant for the kernel AUTO would otherwise have picked.
      if (wpb && o->kernel != MG_KERNEL_AUTO)
        return fail(MG_ERR_INVALID_ARG, "split kernel blocks hold 2, 4 or 8 waves (got %u)", wpb);
      wpb = (G >= 3 || groups * pl->col_splits >= 2) &&
                    split_lds_bytes(G, pl->kl, 4, 2) <= dev->max_lds
                ? 8
                : 2;
    }
    pl->waves_per_block = wpb;
    const int NP = (int)wpb / 2;
    // One group of two or four keys over rows of up to 64 chunks: two DFS waves per pair,
    // each with half the keys (78,643,200 x 1 KiB, one process, two placements: 4 keys
    // 12.55 / 12.74 against 13.22 / 13.35 ms with one DFS wave, 2 keys 11.64 / 12.05 against
    // 11.92 / 12.16 ms). With several groups one DFS wave stays faster (8 keys as 4 + 4:
    // 23.29 against 24.92 ms).
    pl->dfs_waves = (S == 1 && (G == 2 || G == 4) && groups == 1 && pl->col_splits == 1) ? 2 : 1;
    // Mask ring: as many steps as the block's LDS holds, up to kSplitDefaultDepth, so a DFS
    // wave can run ahead through the cheap steps of a tile and cover the steps that descend
    // a whole subtree.
    uint32_t depth = o->mask_depth;
    if (!depth) {
      depth = kSplitDefaultDepth;
      while (depth > 2 && split_lds_bytes(G, pl->kl, NP, depth) > dev->max_lds) depth--;
    }
    pl->mask_depth = depth;
    const size_t lds = split_lds_bytes(G, pl->kl, NP, depth);
    if (lds > dev->max_lds)
      return fail(MG_ERR_INVALID_ARG, "split plan needs %zu B of LDS per block (max %zu)", lds,
                  dev->max_lds);
    int res_waves = 0;
    rc = resident_waves(dev, MG_KERNEL_STREAM_SPLIT * 100 + (uint32_t)(S * 8 + G), S, G,
                        (int)pl->dfs_waves, (int)pl->chunks, (int)wpb, lds, &res_waves);
    // Three pairs per CU (768 on 256 CUs) stream fastest with single-pair blocks; a fourth
    // block on some CUs unbalances them (four keys: 14.9 ms at 768 blocks, 26.5 ms at 832).
    // Four-pair blocks: one per CU.
    resident = std::min(res_waves / 2, kSplitPairsPerCu[NP] * dev->cus);  // pairs
  }
  if (rc) return rc;
  uint64_t parts = o->waves_per_key;
  if (!parts && kernel == MG_KERNEL_LANE_DFS && lane_shape.kl == pl->kl) parts = lane_shape.parts;
  if (!parts) parts = std::max<uint64_t>(1, (uint64_t)resident / (groups * pl->col_splits));
  if (kernel == MG_KERNEL_STREAM_SPLIT || kernel == MG_KERNEL_LANE_DFS) {
    // Waves (lane-DFS) and pairs (split) take equal runs of leaf-pair steps, which may start
    // and end inside a tile, so every resident wave or pair gets work (78,643,200 x 1 KiB at
    // kl 7, 1 key: 9600 tiles are 13 per pair on 739 of 768 pairs, 800 steps on all 768;
    // 11.4-11.7 against 11.7-12.0 ms). The pairs of a split block never wait for each
    // other; lane-DFS blocks meet at a barrier every 2^(kl-1) steps.
    const uint64_t n_pairs = 1ull << (pl->kl - 1);
    const uint64_t total_steps = n_tiles * n_pairs;
    parts = std::min<uint64_t>(parts, total_steps ? total_steps : 1);
    parts = std::min<uint64_t>(parts, 1u << 20);
    pl->steps_per_wave = (total_steps + parts - 1) / parts;
    if (pl->steps_per_wave == 0) pl->steps_per_wave = 1;
    pl->parts = (uint32_t)((total_steps + pl->steps_per_wave - 1) / pl->steps_per_wave);
    // Tiles one pair touches at most.
    pl->tiles_per_wave = (pl->steps_per_wave + n_pairs - 1) / n_pairs + 1;
  } else {
    parts = std::min<uint64_t>(parts, n_tiles ? n_tiles : 1);
    parts = std::min<uint64_t>(parts, 1u << 20);
    pl->tiles_per_wave = (n_tiles + parts - 1) / parts;
    if (pl->tiles_per_wave == 0) pl->tiles_per_wave = 1;
    pl->parts = (uint32_t)((n_tiles + pl->tiles_per_wave - 1) / pl->tiles_per_wave);
  }
  if (pl->parts == 0) pl->parts = 1;
  if ((uint64_t)pl->parts * groups > (1ull << 30))
    return fail(MG_ERR_INVALID_ARG, "%u waves per key x %u key groups is too many", pl->parts,
                groups);
  return MG_OK;
}

int enqueue_scan(mg_device* dev, const Exec& x, const Plan& pl, const ScanBuffers& b,
                 hipEvent_t ev_scan0, hipEvent_t ev_scan1) {
  hipStream_t s = x.stream;
  if (ev_scan0) MG_HIP(hipEventRecord(ev_scan0, s));
  const bool any_rows = pl.row_end > pl.row_begin;
  if (any_rows) {
    if (pl.kernel == MG_KERNEL_LANE_DFS) {
      LaneDfsParams p{};
      p.rows = (const uint4*)pl.rows;
      p.n_rows = pl.row_end;
      p.zero_rows = (const uint4*)dev->zero_rows.p;
      p.keys = b.keys;
      p.key_idx = b.key_idx;
      p.n_keys = pl.n_keys;
      p.d = pl.d; p.kl = pl.kl; p.lb = pl.lb; p.tb = pl.tb;
      p.tile_begin = pl.tile_begin;
      p.tile_end = pl.tile_end;
      p.steps_per_wave = pl.steps_per_wave;
      p.parts = pl.parts;
      p.partial = b.partial;
      p.gstack = b.gstack;
      p.gslots = pl.gslots;
      p.late_loads = pl.late_loads;
      if (pl.gslots && !b.gstack) return fail(MG_ERR_INVALID_ARG, "lane-DFS plan needs a stack buffer");
      if (const uint64_t e = lane_root_entries(pl, pl.n_keys)) {
        // The keys' tile roots first, on the same stream: waves then start each tile at
        // its root instead of walking the tb prefix levels.
        uint4* seeds = b.gstack + lane_wave_region_u4(pl, pl.n_keys);
        uint32_t* ts = reinterpret_cast<uint32_t*>(seeds + e);
        MG_HIP(launch_lane_tile_roots(p, seeds, ts, s));
        p.root_seed = seeds;
        p.root_t = ts;
      }
      MG_HIP(launch_lane_dfs((int)pl.chunks, p, (int)pl.waves_per_block, s));
    } else if (pl.kernel == MG_KERNEL_STREAM || pl.kernel == MG_KERNEL_STREAM_SPLIT) {
      StreamParams p{};
      p.rows = (const uint4*)pl.rows;
      p.n_rows = pl.row_end;
      p.chunks = pl.chunks;
      p.keys = b.keys;
      p.key_idx = b.key_idx;
      p.n_keys = pl.n_keys;
      p.d = pl.d; p.kl = pl.kl; p.lb = pl.lb; p.tb = pl.tb;
      p.tile_begin = pl.tile_begin;
      p.tile_end = pl.tile_end;
      p.tiles_per_wave = pl.tiles_per_wave;
      p.steps_per_wave = pl.steps_per_wave;
      p.parts = pl.parts;
      p.partial = b.partial;
      p.zero_rows = (const uint4*)dev->zero_rows.p;
      if (pl.kernel == MG_KERNEL_STREAM_SPLIT) {
        p.col_splits = pl.col_splits;
        p.mask_depth = pl.mask_depth;
        p.dfs_waves = pl.dfs_waves;
        p.fault = x.fault;
        MG_HIP(launch_split(split_s(&pl), (int)pl.keys_per_wave,
                            (int)pl.waves_per_block / 2, p, s));
      } else {
        MG_HIP(launch_stream(stream_s(pl.chunks), (int)pl.keys_per_wave, stream_pk(pl.chunks), p,
                             (int)pl.waves_per_block, s));
      }
    } else if (pl.kernel == MG_KERNEL_SUFFIX) {
      SuffixParams p{};
      p.rows = (const uint4*)pl.rows;
      p.n_rows = pl.row_end;
      p.keys = b.keys;
      p.key_idx = b.key_idx;
      p.n_keys = pl.n_keys;
      p.d = pl.d;
      p.s = pl.lb;
      p.tile_begin = pl.tile_begin;
      p.n_tiles = pl.tile_end - pl.tile_begin;
      p.partial = b.partial;
      MG_HIP(launch_suffix((int)pl.chunks, p, s));
    } else {
      PointParams p{};
      p.rows = (const uint4*)pl.rows;
      p.row_idx = nullptr;
      p.first_row = pl.row_begin;
      p.n_entries = pl.row_end - pl.row_begin;
      p.chunks = pl.chunks;
      p.keys = b.keys;
      p.key_idx = b.key_idx;
      p.n_keys = pl.n_keys;
      p.partial = b.partial;
      MG_HIP(launch_point(p, s));
    }
  }
  if (ev_scan1) MG_HIP(hipEventRecord(ev_scan1, s));
  // With no rows the reduction over zero parts writes zeros (or leaves `out`
  // unchanged when accumulating).
  MG_HIP(launch_reduce(b.partial, any_rows ? pl.parts : 0, pl.n_keys, pl.chunks, b.out, b.out_idx,
                       b.accumulate, s));
  if (pl.n_delta) {
    PointParams p{};
    p.rows = (const uint4*)pl.delta_rows;
    p.row_idx = (const uint64_t*)pl.delta_idx;
    p.n_entries = pl.n_delta;
    p.chunks = pl.chunks;
    p.keys = b.keys;
    p.key_idx = b.key_idx;
    p.n_keys = pl.n_keys;
    p.partial = b.partial_delta;
    MG_HIP(launch_point(p, s));
    MG_HIP(launch_reduce(b.partial_delta, pl.delta_parts, pl.n_keys, pl.chunks, b.out, b.out_idx,
                         true, s));
  }
  return MG_OK;
}

int attach_delta(mg_device* dev, const mg_matrix* m, const mg_scan_opts* opts, Plan* pl) {
  const mg_delta* dl = opts ? opts->delta : nullptr;
  if (!dl) return MG_OK;
  if (dl->row_bytes != m->row_bytes)
    return fail(MG_ERR_INVALID_ARG, "delta rows are %u B, matrix rows %u B", dl->row_bytes,
                m->row_bytes);
  if (dl->dev != dev) return fail(MG_ERR_INVALID_ARG, "delta belongs to another device");
  if (dl->n && dl->max_idx >= m->n_rows)
    return fail(MG_ERR_ROWS, "delta row %llu outside the %llu-row matrix",
                (unsigned long long)dl->max_idx, (unsigned long long)m->n_rows);
  pl->n_delta = dl->n;
  pl->delta_idx = dl->d_idx;
  pl->delta_rows = dl->d_diffs;
  pl->delta_parts = (uint32_t)((dl->n + 63) / 64);
  return MG_OK;
}

int check_fault(const Exec& x, const Plan& pl) {
  if (pl.kernel != MG_KERNEL_STREAM_SPLIT) return MG_OK;
  uint32_t word = 0;
  MG_HIP(hipMemcpyAsync(&word, x.fault, sizeof word, hipMemcpyDeviceToHost, x.stream));
  MG_HIP(hipStreamSynchronize(x.stream));
  if (!word) return MG_OK;
  const uint32_t zero = 0;
  MG_HIP(hipMemcpyAsync(x.fault, &zero, sizeof zero, hipMemcpyHostToDevice, x.stream));
  MG_HIP(hipStreamSynchronize(x.stream));
  return fail(MG_ERR_DEVICE, "split kernel: a DFS/streaming hand-over wait ran out; answers "
                             "of this call are void");
}

int fill_timing(const hipEvent_t ev[5], const Plan& pl, double t0, mg_timing* t) {
  float h2d = 0, scan = 0, kern = 0, d2h = 0;
  MG_HIP(hipEventElapsedTime(&h2d, ev[0], ev[1]));
  MG_HIP(hipEventElapsedTime(&scan, ev[1], ev[2]));
  MG_HIP(hipEventElapsedTime(&kern, ev[1], ev[3]));
  MG_HIP(hipEventElapsedTime(&d2h, ev[3], ev[4]));
  t->h2d_ms = h2d;
  t->scan_kernel_ms = scan;
  t->kernel_ms = kern;
  t->d2h_ms = d2h;
  t->total_ms = now_ms() - t0;
  t->kernel = pl.kernel;
  t->lane_depth = pl.kl;
  t->stack_global = pl.gslots;
  t->dfs_waves = pl.kernel == MG_KERNEL_STREAM_SPLIT ? pl.dfs_waves : 0;
  const uint32_t n_keys = pl.n_keys;
  const bool stream = pl.kernel == MG_KERNEL_STREAM || pl.kernel == MG_KERNEL_STREAM_SPLIT;
  t->keys_per_wave = stream ? pl.keys_per_wave : 1;
  t->waves = (pl.kernel == MG_KERNEL_POINT || pl.kernel == MG_KERNEL_SUFFIX)
                 ? pl.parts * n_keys
                 : pl.parts * (stream ? (n_keys + pl.keys_per_wave - 1) / pl.keys_per_wave
                                      : n_keys) *
                       (pl.kernel == MG_KERNEL_STREAM_SPLIT ? 2 * pl.col_splits : 1);
  return MG_OK;
}

}  // namespace mg

using mg::DevKey;
using mg::fail;

namespace {

mg::ScanBuffers workspace_buffers(const mg::Workspace* w) {
  mg::ScanBuffers b;
  b.keys = (const DevKey*)w->keys.p;
  b.partial = (uint4*)w->partial.p;
  b.partial_delta = (uint4*)w->partial_delta.p;
  b.out = (uint4*)w->out.p;
  b.gstack = (uint4*)w->gstack.p;
  return b;
}

int run_plan_graph(mg_device* dev, mg::Workspace* w, const mg::Plan& pl) {
  if (!(w->graph_valid && w->graph_plan == pl)) {
    w->invalidate_graph();
    MG_HIP(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
    int rc = mg::enqueue_scan(dev, mg::ws_exec(w), pl, workspace_buffers(w), nullptr, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(w->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return mg::hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&w->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      w->graph_exec = nullptr;
      return mg::hip_fail(e, "hipGraphInstantiate");
    }
    w->graph_plan = pl;
    w->graph_valid = true;
  }
  MG_HIP(hipGraphLaunch(w->graph_exec, w->stream));
  return MG_OK;
}

// A scan enqueued on a workspace, waiting for scan_finish.
struct PendingScan {
  mg::Plan pl;
  double t0 = 0;
  size_t out_bytes = 0;
  uint8_t* out = nullptr;
};

// Parses n_keys wire keys into the workspace's pinned staging.
int stage_keys(mg::Workspace* w, const uint8_t* keys, const size_t* key_lens, uint32_t n_keys,
               uint32_t* d) {
  int rc = w->h_keys.reserve(sizeof(DevKey) * std::max<uint32_t>(n_keys, 1));
  if (rc) return rc;
  return mg::parse_keys(keys, key_lens, n_keys, (DevKey*)w->h_keys.p, d);
}

// Enqueues one scan of the n_keys keys staged in w->h_keys on w's stream: key upload,
// kernels, answer download into w->h_out. Nothing waits for it.
int scan_enqueue(mg_device* dev, mg::Workspace* w, mg_matrix* m, uint32_t n_keys, uint32_t d,
                 uint8_t* out, const mg_scan_opts* opts, PendingScan* ps) {
  ps->t0 = mg::now_ms();
  mg::Plan& pl = ps->pl;
  pl = mg::Plan();
  int rc = mg::make_plan(dev, m->rows, m->n_rows, m->row_bytes, n_keys, d, opts, &pl);
  if (rc) return rc;
  if ((rc = mg::attach_delta(dev, m, opts, &pl))) return rc;
  const size_t out_bytes = (size_t)n_keys * m->row_bytes;
  if ((rc = w->keys.reserve(sizeof(DevKey) * n_keys))) return rc;
  if ((rc = w->out.reserve(out_bytes))) return rc;
  if ((rc = w->partial.reserve((size_t)pl.parts * n_keys * m->row_bytes))) return rc;
  if (pl.gslots && (rc = w->gstack.reserve(mg::lane_gstack_bytes(pl, n_keys)))) return rc;
  if (pl.n_delta &&
      (rc = w->partial_delta.reserve((size_t)pl.delta_parts * n_keys * m->row_bytes)))
    return rc;
  if ((rc = w->h_out.reserve(out_bytes))) return rc;
  pl.ws_keys = w->keys.p;
  pl.ws_out = w->out.p;
  pl.ws_partial = w->partial.p;
  pl.ws_partial_delta = pl.n_delta ? w->partial_delta.p : nullptr;
  pl.ws_gstack = pl.gslots ? w->gstack.p : nullptr;

  hipStream_t s = w->stream;
  MG_HIP(hipEventRecord(w->ev[0], s));
  MG_HIP(hipMemcpyAsync(w->keys.p, w->h_keys.p, sizeof(DevKey) * n_keys,
                        hipMemcpyHostToDevice, s));
  if (opts && opts->use_graph) {
    MG_HIP(hipEventRecord(w->ev[1], s));
    if ((rc = run_plan_graph(dev, w, pl))) return rc;
    MG_HIP(hipEventRecord(w->ev[2], s));
  } else {
    rc = mg::enqueue_scan(dev, mg::ws_exec(w), pl, workspace_buffers(w), w->ev[1], w->ev[2]);
    if (rc) return rc;
  }
  MG_HIP(hipEventRecord(w->ev[3], s));
  MG_HIP(hipMemcpyAsync(w->h_out.p, w->out.p, out_bytes, hipMemcpyDeviceToHost, s));
  MG_HIP(hipEventRecord(w->ev[4], s));
  ps->out_bytes = out_bytes;
  ps->out = out;
  return MG_OK;
}

// Waits for a scan_enqueue'd scan and copies its answers out.
int scan_finish(mg::Workspace* w, const PendingScan& ps, mg_timing* t) {
  MG_HIP(hipStreamSynchronize(w->stream));
  int rc = mg::check_fault(mg::ws_exec(w), ps.pl);
  if (rc) return rc;
  memcpy(ps.out, w->h_out.p, ps.out_bytes);
  if (t) return mg::fill_timing(w->ev, ps.pl, ps.t0, t);
  return MG_OK;
}

}  // namespace

// ------------------------------------------------------------------- ABI --

extern "C" {

int mg_abi_version(void) { return MG_ABI_VERSION; }

const char* mg_last_error(void) { return mg::last_error(); }

int mg_device_open(int ordinal, mg_device** out) {
  if (!out) return fail(MG_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(MG_ERR_NO_DEVICE, "no HIP device available");
  if (ordinal < 0 || ordinal >= count)
    return fail(MG_ERR_NO_DEVICE, "device %d not present (%d devices)", ordinal, count);
  MG_HIP(hipSetDevice(ordinal));
  hipDeviceProp_t prop;
  MG_HIP(hipGetDeviceProperties(&prop, ordinal));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(MG_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", ordinal,
                prop.gcnArchName);
  mg_device* dev = new mg_device();
  dev->ordinal = ordinal;
  dev->cus = prop.multiProcessorCount;
  dev->max_lds = prop.sharedMemPerBlock;
  hipError_t se = hipStreamCreateWithFlags(&dev->stream, hipStreamNonBlocking);
  if (se != hipSuccess) {
    delete dev;
    return mg::hip_fail(se, "hipStreamCreateWithFlags");
  }
  for (auto& ev : dev->ev) {
    hipError_t ee = hipEventCreate(&ev);
    if (ee != hipSuccess) {
      mg_device_close(dev);
      return mg::hip_fail(ee, "hipEventCreate");
    }
  }
  int rc = dev->zero_rows.reserve(4096);
  if (!rc) rc = dev->fault.reserve(16);
  if (!rc && (mg::launch_zero((uint4*)dev->zero_rows.p, 4096 / 16, dev->stream) != hipSuccess ||
              mg::launch_zero((uint4*)dev->fault.p, 1, dev->stream) != hipSuccess ||
              hipStreamSynchronize(dev->stream) != hipSuccess))
    rc = fail(MG_ERR_DEVICE, "zeroing the zero-row buffer failed");
  if (!rc) rc = mg::noalloc_setup(dev);
  if (rc) {
    mg_device_close(dev);
    return rc;
  }
  *out = dev;
  return MG_OK;
}

int mg_device_close(mg_device* dev) {
  if (!dev) return MG_OK;
  if (const int n = dev->children.load())
    return fail(MG_ERR_STATE, "%d matrices, deltas or request loops still use this device", n);
  (void)hipSetDevice(dev->ordinal);
  if (dev->stream) (void)hipStreamSynchronize(dev->stream);
  for (mg::Workspace& w : dev->ws) {
    std::lock_guard<std::mutex> lk(w.mu);
    w.release();
  }
  dev->zero_rows.release();
  dev->fault.release();
  for (auto& ev : dev->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (dev->stream) (void)hipStreamDestroy(dev->stream);
  delete dev;
  return MG_OK;
}

int mg_device_set_workspaces(mg_device* dev, uint32_t n) {
  if (!dev) return fail(MG_ERR_INVALID_ARG, "device is null");
  if (n < 1 || n > mg::kMaxWorkspaces)
    return fail(MG_ERR_INVALID_ARG, "%u workspaces: 1 to %u", n, mg::kMaxWorkspaces);
  dev->n_workspaces = n;
  return MG_OK;
}

int mg_device_info(mg_device* dev, uint64_t* free_bytes, uint64_t* total_bytes,
                   uint32_t* compute_units, char* name, size_t name_len) {
  if (!dev) return fail(MG_ERR_INVALID_ARG, "device is null");
  int rc = mg::set_device(dev);
  if (rc) return rc;
  size_t fr = 0, tot = 0;
  MG_HIP(hipMemGetInfo(&fr, &tot));
  if (free_bytes) *free_bytes = fr;
  if (total_bytes) *total_bytes = tot;
  if (compute_units) *compute_units = (uint32_t)dev->cus;
  if (name && name_len) {
    hipDeviceProp_t prop;
    MG_HIP(hipGetDeviceProperties(&prop, dev->ordinal));
    snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  return MG_OK;
}

int mg_matrix_create(mg_device* dev, uint64_t n_rows, uint32_t row_bytes, mg_matrix** out) {
  if (!dev || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (row_bytes == 0 || row_bytes % 16 != 0)
    return fail(MG_ERR_ROWS, "row width %u is not a positive multiple of 16", row_bytes);
  if (n_rows > (1ull << 32)) return fail(MG_ERR_ROWS, "more than 2^32 rows");
  std::lock_guard<std::mutex> lk(dev->mu);
  int rc = mg::set_device(dev);
  if (rc) return rc;
  mg_matrix* m = new mg_matrix();
  m->dev = dev;
  m->n_rows = n_rows;
  m->row_bytes = row_bytes;
  const size_t bytes = std::max<size_t>((size_t)n_rows * row_bytes, 16);
  hipError_t e = hipMalloc((void**)&m->rows, bytes);
  if (e != hipSuccess) {
    delete m;
    return mg::hip_fail(e, "hipMalloc(matrix)");
  }
  if (mg::launch_zero(m->rows, bytes / 16, dev->stream) != hipSuccess ||
      hipStreamSynchronize(dev->stream) != hipSuccess) {
    (void)hipFree(m->rows);
    delete m;
    return fail(MG_ERR_DEVICE, "zeroing the matrix failed");
  }
  dev->children++;
  *out = m;
  return MG_OK;
}

int mg_matrix_upload(mg_matrix* m, uint64_t first_row, const void* rows, uint64_t n_rows) {
  if (!m || (!rows && n_rows)) return fail(MG_ERR_INVALID_ARG, "null argument");
  if (first_row > m->n_rows || n_rows > m->n_rows - first_row)
    return fail(MG_ERR_ROWS, "rows [%llu, +%llu) exceed the %llu-row matrix",
                (unsigned long long)first_row, (unsigned long long)n_rows,
                (unsigned long long)m->n_rows);
  std::lock_guard<std::mutex> lk(m->dev->mu);
  std::unique_lock<mg::ScanGate> wlk(m->dev->state);
  int rc = mg::set_device(m->dev);
  if (rc) return rc;
  MG_HIP(hipMemcpyAsync((uint8_t*)m->rows + first_row * m->row_bytes, rows,
                        (size_t)n_rows * m->row_bytes, hipMemcpyHostToDevice, m->dev->stream));
  MG_HIP(hipStreamSynchronize(m->dev->stream));
  return MG_OK;
}

int mg_matrix_download(mg_matrix* m, uint64_t first_row, void* rows, uint64_t n_rows) {
  if (!m || (!rows && n_rows)) return fail(MG_ERR_INVALID_ARG, "null argument");
  if (first_row > m->n_rows || n_rows > m->n_rows - first_row)
    return fail(MG_ERR_ROWS, "rows [%llu, +%llu) exceed the %llu-row matrix",
                (unsigned long long)first_row, (unsigned long long)n_rows,
                (unsigned long long)m->n_rows);
  std::lock_guard<std::mutex> lk(m->dev->mu);
  int rc = mg::set_device(m->dev);
  if (rc) return rc;
  MG_HIP(hipMemcpyAsync(rows, (const uint8_t*)m->rows + first_row * m->row_bytes,
                        (size_t)n_rows * m->row_bytes, hipMemcpyDeviceToHost, m->dev->stream));
  MG_HIP(hipStreamSynchronize(m->dev->stream));
  return MG_OK;
}

int mg_matrix_fill_seeded(mg_matrix* m, uint64_t seed) {
  if (!m) return fail(MG_ERR_INVALID_ARG, "matrix is null");
  std::lock_guard<std::mutex> lk(m->dev->mu);
  std::unique_lock<mg::ScanGate> wlk(m->dev->state);
  int rc = mg::set_device(m->dev);
  if (rc) return rc;
  const uint64_t pairs = (uint64_t)m->n_rows * m->row_bytes / 16;
  if (pairs) MG_HIP(mg::launch_fill_splitmix(m->rows, 0, pairs, seed, m->dev->stream));
  MG_HIP(hipStreamSynchronize(m->dev->stream));
  return MG_OK;
}

int mg_matrix_fill_pattern(mg_matrix* m, uint64_t seed) {
  if (!m) return fail(MG_ERR_INVALID_ARG, "matrix is null");
  std::lock_guard<std::mutex> lk(m->dev->mu);
  std::unique_lock<mg::ScanGate> wlk(m->dev->state);
  int rc = mg::set_device(m->dev);
  if (rc) return rc;
  const uint64_t blocks = (uint64_t)m->n_rows * m->row_bytes / 16;
  if (blocks) MG_HIP(mg::launch_fill_lcg(m->rows, blocks, seed, m->dev->stream));
  MG_HIP(hipStreamSynchronize(m->dev->stream));
  return MG_OK;
}

int mg_matrix_info(const mg_matrix* m, uint64_t* n_rows, uint32_t* row_bytes) {
  if (!m) return fail(MG_ERR_INVALID_ARG, "matrix is null");
  if (n_rows) *n_rows = m->n_rows;
  if (row_bytes) *row_bytes = m->row_bytes;
  return MG_OK;
}

int mg_matrix_copy(mg_matrix* dst, const mg_matrix* src) {
  if (!dst || !src) return fail(MG_ERR_INVALID_ARG, "null argument");
  if (dst->dev != src->dev) return fail(MG_ERR_INVALID_ARG, "matrices on different devices");
  if (dst->n_rows != src->n_rows || dst->row_bytes != src->row_bytes)
    return fail(MG_ERR_ROWS, "matrix shapes differ");
  std::lock_guard<std::mutex> lk(dst->dev->mu);
  std::unique_lock<mg::ScanGate> wlk(dst->dev->state);
  int rc = mg::set_device(dst->dev);
  if (rc) return rc;
  const uint64_t n = dst->n_rows * dst->row_bytes / 16;
  if (n && dst != src) MG_HIP(mg::launch_copy(dst->rows, src->rows, n, dst->dev->stream));
  MG_HIP(hipStreamSynchronize(dst->dev->stream));
  return MG_OK;
}

int mg_matrix_destroy(mg_matrix* m) {
  if (!m) return MG_OK;
  if (const int n = m->users.load())
    return fail(MG_ERR_STATE, "%d request loops still serve this matrix", n);
  mg_device* dev = m->dev;
  std::lock_guard<std::mutex> lk(dev->mu);
  std::unique_lock<mg::ScanGate> wlk(dev->state);
  (void)mg::set_device(dev);
  (void)hipStreamSynchronize(dev->stream);
  mg::invalidate_graphs(dev);
  (void)hipFree(m->rows);
  delete m;
  dev->children--;
  return MG_OK;
}

int mg_key_parse(const uint8_t* bytes, size_t len, mg_key_info* out) {
  if (!bytes || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
  DevKey k;
  uint32_t party = 0;
  switch (mg::parse_wire_key(bytes, len, &k, &party)) {
    case mg::kParseOk:
      break;
    case mg::kParseFormat:
      return fail(MG_ERR_KEY_FORMAT, "invalid key format (length %zu)", len);
    default:
      return fail(MG_ERR_DOMAIN, "domain bits %u > %d", (unsigned)bytes[18], mg::kMaxDomainBits);
  }
  out->domain_bits = k.domain_bits;
  out->party = party;
  out->root_t = k.root_t;
  out->wire_len = (uint32_t)len;
  return MG_OK;
}

int mg_scan(mg_device* dev, mg_matrix* m, const uint8_t* keys, const size_t* key_lens,
            uint32_t n_keys, uint8_t* out, const mg_scan_opts* opts, mg_timing* timing) {
  if (!dev || !m) return fail(MG_ERR_INVALID_ARG, "null device or matrix");
  if (m->dev != dev) return fail(MG_ERR_INVALID_ARG, "matrix belongs to another device");
  if (n_keys == 0) return MG_OK;
  if (!keys || !key_lens || !out) return fail(MG_ERR_INVALID_ARG, "null keys or output");
  std::shared_lock<mg::ScanGate> slk(dev->state);
  int rc = mg::set_device(dev);
  if (rc) return rc;
  mg::WorkspaceLease w(dev);
  if ((rc = w->ready())) return rc;
  uint32_t d = 0;
  if ((rc = stage_keys(w.get(), keys, key_lens, n_keys, &d))) return rc;
  PendingScan ps;
  rc = scan_enqueue(dev, w.get(), m, n_keys, d, out, opts, &ps);
  if (rc) {
    (void)hipStreamSynchronize(w->stream);  // nothing may still use the staging buffers
    return rc;
  }
  return scan_finish(w.get(), ps, timing);
}

int mg_scan_per_query(mg_device* dev, mg_matrix* m, const uint8_t* keys, const size_t* key_lens,
                      uint32_t n_keys, uint32_t keys_per_query, uint8_t* out,
                      const mg_scan_opts* opts, mg_timing* timing) {
  if (keys_per_query == 0) return fail(MG_ERR_INVALID_ARG, "keys_per_query is 0");
  if (!dev || !m) return fail(MG_ERR_INVALID_ARG, "null device or matrix");
  if (m->dev != dev) return fail(MG_ERR_INVALID_ARG, "matrix belongs to another device");
  if (n_keys == 0) return MG_OK;
  if (!keys || !key_lens || !out) return fail(MG_ERR_INVALID_ARG, "null keys or output");
  std::shared_lock<mg::ScanGate> slk(dev->state);
  int rc = mg::set_device(dev);
  if (rc) return rc;
  // One launch sequence per query, dealt round-robin over the device's first k
  // workspaces (taken in index order, so callers doing the same cannot deadlock), each
  // on its own stream: query q + k is enqueued once query q's answers are out.
  const uint32_t n_q = (n_keys + keys_per_query - 1) / keys_per_query;
  const uint32_t k = std::min(std::min(dev->n_workspaces.load(), mg::kMaxWorkspaces), n_q);
  std::vector<std::unique_ptr<mg::WorkspaceLease>> ws;
  for (uint32_t j = 0; j < k; j++) {
    ws.emplace_back(new mg::WorkspaceLease(dev, j));
    if ((rc = (*ws.back())->ready())) return rc;
  }
  std::vector<PendingScan> pend(k);
  std::vector<bool> busy(k, false);
  mg_timing acc{};
  const double t0 = mg::now_ms();
  auto finish = [&](uint32_t j) -> int {
    busy[j] = false;
    mg_timing t{};
    int r = scan_finish(ws[j]->get(), pend[j], &t);
    if (r) return r;
    acc.h2d_ms += t.h2d_ms;
    acc.kernel_ms += t.kernel_ms;
    acc.d2h_ms += t.d2h_ms;
    acc.scan_kernel_ms += t.scan_kernel_ms;
    acc.kernel = t.kernel;
    acc.lane_depth = t.lane_depth;
    acc.keys_per_wave = t.keys_per_wave;
    acc.stack_global = t.stack_global;
    acc.waves = std::max(acc.waves, t.waves);
    return MG_OK;
  };
  size_t off = 0;
  for (uint32_t q = 0, qi = 0; q < n_keys && !rc; q += keys_per_query, qi++) {
    const uint32_t j = qi % k;
    const uint32_t nk = std::min(keys_per_query, n_keys - q);
    mg::Workspace* w = ws[j]->get();
    if (busy[j] && (rc = finish(j))) break;
    uint32_t d = 0;
    if ((rc = stage_keys(w, keys + off, key_lens + q, nk, &d))) break;
    for (uint32_t i = 0; i < nk; i++) off += key_lens[q + i];
    rc = scan_enqueue(dev, w, m, nk, d, out + (size_t)q * m->row_bytes, opts, &pend[j]);
    busy[j] = true;
  }
  for (uint32_t j = 0; j < k; j++) {
    if (!busy[j]) continue;
    if (rc)
      (void)hipStreamSynchronize((*ws[j])->stream);  // drain before the staging is reused
    else
      rc = finish(j);
  }
  if (rc) return rc;
  acc.total_ms = mg::now_ms() - t0;
  if (timing) *timing = acc;
  return MG_OK;
}

int mg_delta_create(mg_device* dev, uint32_t row_bytes, uint64_t capacity, mg_delta** out) {
  if (!dev || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (row_bytes == 0 || row_bytes % 16 != 0)
    return fail(MG_ERR_ROWS, "row width %u is not a positive multiple of 16", row_bytes);
  std::lock_guard<std::mutex> lk(dev->mu);
  int rc = mg::set_device(dev);
  if (rc) return rc;
  mg_delta* dl = new mg_delta();
  dl->dev = dev;
  dl->row_bytes = row_bytes;
  dl->capacity = capacity;
  const size_t cap = std::max<uint64_t>(capacity, 1);
  hipError_t e = hipMalloc((void**)&dl->d_idx, cap * 8);
  if (e == hipSuccess) e = hipMalloc((void**)&dl->d_diffs, cap * row_bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&dl->d_uidx, cap * 8);
  if (e == hipSuccess) e = hipMalloc((void**)&dl->d_udiffs, cap * row_bytes);
  if (e != hipSuccess) {
    (void)hipFree(dl->d_idx);
    (void)hipFree(dl->d_diffs);
    (void)hipFree(dl->d_uidx);
    (void)hipFree(dl->d_udiffs);
    delete dl;
    return mg::hip_fail(e, "hipMalloc(delta)");
  }
  dev->children++;
  *out = dl;
  return MG_OK;
}

int mg_delta_set(mg_delta* dl, const uint64_t* row_idx, const uint8_t* diffs, uint64_t n_entries,
                 uint64_t epoch) {
  if (!dl || (n_entries && (!row_idx || !diffs))) return fail(MG_ERR_INVALID_ARG, "null argument");
  if (n_entries > dl->capacity)
    return fail(MG_ERR_INVALID_ARG, "%llu entries exceed the delta capacity %llu",
                (unsigned long long)n_entries, (unsigned long long)dl->capacity);
  uint64_t mx = 0;
  for (uint64_t i = 0; i < n_entries; i++) mx = std::max(mx, row_idx[i]);
  std::vector<uint64_t> uidx;
  std::vector<uint8_t> udiff;
  mg::dedupe_delta(row_idx, diffs, n_entries, dl->row_bytes, &uidx, &udiff);
  std::lock_guard<std::mutex> lk(dl->dev->mu);
  std::unique_lock<mg::ScanGate> wlk(dl->dev->state);
  int rc = mg::set_device(dl->dev);
  if (rc) return rc;
  hipStream_t s = dl->dev->stream;
  if (n_entries) {
    MG_HIP(hipMemcpyAsync(dl->d_idx, row_idx, n_entries * 8, hipMemcpyHostToDevice, s));
    MG_HIP(hipMemcpyAsync(dl->d_diffs, diffs, n_entries * dl->row_bytes, hipMemcpyHostToDevice, s));
    MG_HIP(hipMemcpyAsync(dl->d_uidx, uidx.data(), uidx.size() * 8, hipMemcpyHostToDevice, s));
    MG_HIP(hipMemcpyAsync(dl->d_udiffs, udiff.data(), udiff.size(), hipMemcpyHostToDevice, s));
  }
  MG_HIP(hipStreamSynchronize(s));
  dl->n = n_entries;
  dl->n_unique = uidx.size();
  dl->max_idx = mx;
  dl->epoch = epoch;
  return MG_OK;
}

int mg_delta_info(const mg_delta* dl, uint64_t* n_entries, uint64_t* epoch) {
  if (!dl) return fail(MG_ERR_INVALID_ARG, "delta is null");
  if (n_entries) *n_entries = dl->n;
  if (epoch) *epoch = dl->epoch;
  return MG_OK;
}

int mg_delta_destroy(mg_delta* dl) {
  if (!dl) return MG_OK;
  mg_device* dev = dl->dev;
  std::lock_guard<std::mutex> lk(dev->mu);
  std::unique_lock<mg::ScanGate> wlk(dev->state);
  (void)mg::set_device(dev);
  (void)hipStreamSynchronize(dev->stream);
  mg::invalidate_graphs(dev);
  (void)hipFree(dl->d_idx);
  (void)hipFree(dl->d_diffs);
  (void)hipFree(dl->d_uidx);
  (void)hipFree(dl->d_udiffs);
  delete dl;
  dev->children--;
  return MG_OK;
}

int mg_scan_delta(mg_device* dev, const mg_delta* dl, const uint8_t* keys, const size_t* key_lens,
                  uint32_t n_keys, uint8_t* out) {
  if (!dev || !dl) return fail(MG_ERR_INVALID_ARG, "null device or delta");
  if (dl->dev != dev) return fail(MG_ERR_INVALID_ARG, "delta belongs to another device");
  if (n_keys == 0) return MG_OK;
  if (!keys || !key_lens || !out) return fail(MG_ERR_INVALID_ARG, "null keys or output");
  std::shared_lock<mg::ScanGate> slk(dev->state);
  int rc = mg::set_device(dev);
  if (rc) return rc;
  mg::WorkspaceLease w(dev);
  if ((rc = w->ready())) return rc;
  uint32_t d = 0;
  if ((rc = stage_keys(w.get(), keys, key_lens, n_keys, &d))) return rc;
  if (d < 64 && dl->n && (dl->max_idx >> d) != 0)
    return fail(MG_ERR_ROWS, "delta row %llu outside the 2^%u-leaf domain",
                (unsigned long long)dl->max_idx, d);
  const uint32_t chunks = dl->row_bytes / 16;
  const size_t out_bytes = (size_t)n_keys * dl->row_bytes;
  const uint32_t parts = (uint32_t)((dl->n + 63) / 64);
  if ((rc = w->keys.reserve(sizeof(DevKey) * n_keys))) return rc;
  if ((rc = w->out.reserve(out_bytes))) return rc;
  if ((rc = w->partial_delta.reserve((size_t)std::max(parts, 1u) * out_bytes))) return rc;
  if ((rc = w->h_out.reserve(out_bytes))) return rc;
  hipStream_t s = w->stream;
  MG_HIP(hipMemcpyAsync(w->keys.p, w->h_keys.p, sizeof(DevKey) * n_keys,
                        hipMemcpyHostToDevice, s));
  if (dl->n) {
    mg::PointParams p{};
    p.rows = dl->d_diffs;
    p.row_idx = dl->d_idx;
    p.n_entries = dl->n;
    p.chunks = chunks;
    p.keys = (const DevKey*)w->keys.p;
    p.n_keys = n_keys;
    p.partial = (uint4*)w->partial_delta.p;
    MG_HIP(mg::launch_point(p, s));
  }
  MG_HIP(mg::launch_reduce((const uint4*)w->partial_delta.p, parts, n_keys, chunks,
                           (uint4*)w->out.p, nullptr, false, s));
  MG_HIP(hipMemcpyAsync(w->h_out.p, w->out.p, out_bytes, hipMemcpyDeviceToHost, s));
  MG_HIP(hipStreamSynchronize(s));
  memcpy(out, w->h_out.p, out_bytes);
  return MG_OK;
}

int mg_matrix_apply_delta(mg_matrix* m, const mg_delta* dl) {
  if (!m || !dl) return fail(MG_ERR_INVALID_ARG, "null argument");
  if (dl->dev != m->dev) return fail(MG_ERR_INVALID_ARG, "delta belongs to another device");
  if (dl->row_bytes != m->row_bytes)
    return fail(MG_ERR_INVALID_ARG, "delta rows are %u B, matrix rows %u B", dl->row_bytes,
                m->row_bytes);
  if (dl->n && dl->max_idx >= m->n_rows)
    return fail(MG_ERR_ROWS, "delta row %llu outside the %llu-row matrix",
                (unsigned long long)dl->max_idx, (unsigned long long)m->n_rows);
  std::lock_guard<std::mutex> lk(m->dev->mu);
  std::unique_lock<mg::ScanGate> wlk(m->dev->state);
  int rc = mg::set_device(m->dev);
  if (rc) return rc;
  if (dl->n_unique)
    MG_HIP(mg::launch_xor_rows(m->rows, dl->d_uidx, dl->d_udiffs, dl->n_unique,
                               m->row_bytes / 16, m->dev->stream));
  MG_HIP(hipStreamSynchronize(m->dev->stream));
  return MG_OK;
}

}  // extern "C"
