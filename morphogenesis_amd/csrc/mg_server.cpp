// Continuous-batching request loop over one device-resident matrix.
//
// Replaces the reference server's GPU batch policy and dispatch
// (crates/morphogen-server/src/network/gpu_batch.rs:129-149, gpu_scan.rs:36-76)
// and its epoch-consistent scan wrapper (scan.rs:260-299) with a rolling sweep:
//  - The matrix is cut into S segments of whole tiles. One sweep step runs the
//    fused scan over one segment for every active key, XOR-accumulating into
//    the key's device-resident accumulator.
//  - A request is admitted, FIFO, as soon as enough key slots are free; its
//    keys join at the next step and finish after S steps. The XOR fold is
//    order-free, so each answer equals a whole-matrix scan.
//  - Updates (epochs) append to a cumulative delta over the current base
//    rows, at the next sweep step, ahead of requests still waiting for slots
//    (the reference's merge worker advances epochs whatever queries wait,
//    epoch.rs:465-549, 685). A key's accumulator starts at the delta
//    contribution of every update applied before its admission
//    (scan_delta_for_gpu, scan.rs:434-485), so its answer is the scan of base
//    XOR updates at admission, however the sweep and later updates interleave
//    with it; the request reports that epoch.
//  - Overload: with a deadline, a request whose predicted completion (work
//    queued and in flight ahead of it, times the measured GPU time per key
//    and sweep step) exceeds it is refused at submit (MG_ERR_BUSY, the
//    reference's 503, network/api/mod.rs:638, 828). An abandoned request
//    leaves the queue, or the sweep at the next step, so a client that gave
//    up costs no more GPU time.
//  - When the cumulative delta grows past half its capacity the loop merges it
//    into a base copy: into the spare copy when double-buffered (in-flight
//    keys keep the old base), otherwise in place once in-flight keys drain.
//  - A snapshot (EpochManager::submit_snapshot, epoch.rs:407-463) replaces
//    the served matrix: requests admitted before it finish on the old rows,
//    the pending updates are dropped (the new rows include them), and later
//    requests start on the new rows at the snapshot's epoch.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>

#include "mg_internal.hpp"

using mg::DevKey;
using mg::fail;

namespace {

struct Request {
  uint64_t ticket = 0;
  uint32_t n_keys = 0;
  std::vector<DevKey> keys;
  std::vector<uint8_t> answers;
  uint32_t keys_done = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  uint64_t epoch = 0;
  int status = MG_OK;
  std::string error;
  bool done = false;
  bool admitted = false;
  bool abandoned = false;  // nobody will collect it: dropped once done
  bool cancel = false;     // abandoned in flight: its keys leave at the next step
  uint32_t keys_left = 0;  // keys still in slots (admitted requests)
  uint32_t steps_done = 0; // sweep steps its keys have run (they run in lockstep)
  double due = 0;          // admission deadline, absolute ms (0 = none)
  int waiting = 0;         // threads in mg_server_wait for it
};

struct Update {
  std::vector<uint64_t> idx;
  std::vector<uint8_t> diffs;
  uint64_t epoch = 0;
};

// A matrix replacing the served one; owned by the caller blocked in
// mg_server_snapshot until `done`.
struct Snapshot {
  mg_matrix* m = nullptr;
  uint64_t epoch = 0;
  bool done = false;
  int status = MG_OK;
  std::string error;
};

// One entry of the submission-ordered queue: a request, an update or a
// snapshot.
struct Event {
  Request* req = nullptr;
  std::unique_ptr<Update> upd;
  Snapshot* snap = nullptr;
};

struct Slot {
  Request* req = nullptr;
  uint32_t key = 0;        // index of the key within its request
  uint32_t steps_done = 0;
  int base = 0;
};

}  // namespace

struct mg_server {
  mg_device* dev = nullptr;
  mg_matrix* m = nullptr;
  uint32_t row_bytes = 0, chunks = 0, d = 0, kl = 0, segments = 0;
  uint64_t n_rows = 0, seg_rows = 0;
  uint64_t queued_n_rows = 0;  // rows after every queued snapshot (under mu)
  uint32_t seg_want = 0;       // segments option (0 = one per 2^25 rows)
  uint32_t cap_keys = 0;
  bool double_buffer = false;
  uint32_t deadline_ms = 0;
  bool use_graph = false;

  // Base row copies: 0 is the caller's matrix, 1 the spare (double buffer).
  uint4* base_rows[2] = {nullptr, nullptr};
  int cur_base = 0;
  bool rebase_wanted = false;

  // Cumulative updates over the current base (host mirror + device copy).
  std::vector<uint64_t> c_idx;
  std::vector<uint8_t> c_diffs;
  uint64_t c_cap = 0;
  uint64_t* d_cidx = nullptr;
  uint4* d_cdiffs = nullptr;
  uint64_t rebases = 0;
  uint64_t graph_launches = 0;  // loop thread only

  // Key slots, device-resident keys and accumulators.
  std::vector<Slot> slots;
  std::vector<uint32_t> free_slots;
  uint32_t active = 0;
  DevKey* d_keys = nullptr;
  uint4* d_acc = nullptr;
  uint32_t* d_idx = nullptr;  // slot lists: [0, cap) scan list, [cap, 2cap) admission list
  DevKey* d_stage_keys = nullptr;
  mg::DevBuf partial, partial_delta;
  mg::DevBuf gstack;  // lane-DFS stack levels in global memory
  mg::HostBuf h_idx, h_keys, h_acc;
  uint32_t step = 0;
  // Sweep-step graphs, one per (segment, base copy), each for the plan it
  // was captured with.
  struct StepGraph {
    bool valid = false;
    mg::Plan plan;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<StepGraph> graphs;

  // Queue, shared with the API threads under `mu`.
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Event> queue;
  std::map<uint64_t, std::unique_ptr<Request>> requests;
  uint64_t next_ticket = 1;
  uint64_t epoch = 0;
  // Admission model (under mu): requests in flight, and a step's time as
  // step_fixed_ms + key_step_ms x active keys, fitted by exponentially
  // weighted least squares over the steps run so far.
  std::vector<Request*> inflight;
  uint64_t queued_keys = 0;
  double fit_w = 0, fit_n = 0, fit_t = 0, fit_nn = 0, fit_nt = 0;
  double key_step_ms = 0, step_fixed_ms = 0;  // the model in use
  double prior_key_ms = 0, prior_fixed_ms = 0;  // before enough steps are timed
  // Threads inside mg_server_wait / mg_server_snapshot (under mu): destroy
  // frees the server only once they have left.
  int waiters = 0;
  std::condition_variable cv_gone;
  int inject = MG_OK;  // mg_server_inject_failure: the next turn fails with this code
  bool stop = false;
  int fatal = MG_OK;
  std::string fatal_msg;
  mg_server_stats stats{};
  std::thread thread;
};

namespace {

void drop_graphs(mg_server* s) {
  for (auto& g : s->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = mg_server::StepGraph{};
  }
}

void free_server_memory(mg_server* s) {
  drop_graphs(s);
  (void)hipFree(s->d_cidx);
  (void)hipFree(s->d_cdiffs);
  (void)hipFree(s->d_keys);
  (void)hipFree(s->d_acc);
  (void)hipFree(s->d_idx);
  (void)hipFree(s->d_stage_keys);
  if (s->base_rows[1]) (void)hipFree(s->base_rows[1]);
  s->partial.release();
  s->partial_delta.release();
  s->gstack.release();
  s->h_idx.release();
  s->h_keys.release();
  s->h_acc.release();
}

// Keys pinned to base b that are still in flight.
uint32_t keys_on_base(const mg_server* s, int b) {
  uint32_t n = 0;
  for (const Slot& sl : s->slots)
    if (sl.req && sl.base == b) n++;
  return n;
}

// Merges the cumulative updates into a base copy if that is possible now.
// Caller holds dev->mu.
int try_rebase(mg_server* s) {
  if (!s->rebase_wanted) return MG_OK;
  hipStream_t st = s->dev->stream;
  int target;
  if (s->double_buffer) {
    target = 1 - s->cur_base;
    if (keys_on_base(s, target)) return MG_OK;  // old keys still read it
  } else {
    target = s->cur_base;
    if (s->active) return MG_OK;  // merge in place once nothing reads the base
  }
  // Base 0 is the caller's matrix: scans of it on other streams wait while it changes.
  std::unique_lock<mg::ScanGate> wlk(s->dev->state, std::defer_lock);
  if (target == 0) wlk.lock();
  if (s->double_buffer)
    MG_HIP(hipMemcpyAsync(s->base_rows[target], s->base_rows[s->cur_base],
                          (size_t)s->n_rows * s->row_bytes, hipMemcpyDeviceToDevice, st));
  std::vector<uint64_t> uidx;
  std::vector<uint8_t> udiff;
  mg::dedupe_delta(s->c_idx.data(), s->c_diffs.data(), s->c_idx.size(), s->row_bytes, &uidx,
                   &udiff);
  if (!uidx.empty()) {
    // The cumulative-delta buffers are free to reuse as staging: they are cleared below.
    MG_HIP(hipMemcpyAsync(s->d_cidx, uidx.data(), uidx.size() * 8, hipMemcpyHostToDevice, st));
    MG_HIP(hipMemcpyAsync(s->d_cdiffs, udiff.data(), udiff.size(), hipMemcpyHostToDevice, st));
    MG_HIP(mg::launch_xor_rows(s->base_rows[target], s->d_cidx, s->d_cdiffs, uidx.size(),
                               s->chunks, st));
  }
  MG_HIP(hipStreamSynchronize(st));
  s->cur_base = target;
  s->c_idx.clear();
  s->c_diffs.clear();
  s->rebase_wanted = false;
  s->rebases++;
  return MG_OK;
}

// Appends one update to the cumulative delta. Caller holds dev->mu.
int apply_update(mg_server* s, const Update& u) {
  const uint64_t off = s->c_idx.size();
  const uint64_t n = u.idx.size();
  if (n) {
    hipStream_t st = s->dev->stream;
    MG_HIP(hipMemcpyAsync(s->d_cidx + off, u.idx.data(), n * 8, hipMemcpyHostToDevice, st));
    MG_HIP(hipMemcpyAsync((uint8_t*)s->d_cdiffs + off * s->row_bytes, u.diffs.data(),
                          u.diffs.size(), hipMemcpyHostToDevice, st));
    MG_HIP(hipStreamSynchronize(st));
  }
  s->c_idx.insert(s->c_idx.end(), u.idx.begin(), u.idx.end());
  s->c_diffs.insert(s->c_diffs.end(), u.diffs.begin(), u.diffs.end());
  if (s->c_idx.size() * 2 > s->c_cap) s->rebase_wanted = true;
  return MG_OK;
}

// Places the keys of newly admitted requests in free slots and sets their
// accumulators to the contribution of the cumulative delta. Caller holds
// dev->mu.
int admit_keys(mg_server* s, const std::vector<Request*>& reqs) {
  std::vector<uint32_t> new_slots;
  std::vector<DevKey> keys;
  for (Request* r : reqs)
    for (uint32_t k = 0; k < r->n_keys; k++) {
      const uint32_t sl = s->free_slots.back();
      s->free_slots.pop_back();
      s->slots[sl] = Slot{r, k, 0, s->cur_base};
      new_slots.push_back(sl);
      keys.push_back(r->keys[k]);
    }
  const uint32_t n = (uint32_t)new_slots.size();
  if (!n) return MG_OK;
  hipStream_t st = s->dev->stream;
  uint32_t* h_idx = (uint32_t*)s->h_idx.p + s->cap_keys;
  memcpy(h_idx, new_slots.data(), n * 4);
  memcpy(s->h_keys.p, keys.data(), n * sizeof(DevKey));
  uint32_t* d_list = s->d_idx + s->cap_keys;
  MG_HIP(hipMemcpyAsync(d_list, h_idx, n * 4, hipMemcpyHostToDevice, st));
  MG_HIP(hipMemcpyAsync(s->d_stage_keys, s->h_keys.p, n * sizeof(DevKey), hipMemcpyHostToDevice,
                        st));
  MG_HIP(mg::launch_scatter_u4((uint4*)s->d_keys, (const uint4*)s->d_stage_keys, d_list, n,
                               sizeof(DevKey) / 16, st));
  const uint64_t c_n = s->c_idx.size();
  const uint32_t parts = (uint32_t)((c_n + 63) / 64);
  int rc;
  if (c_n) {
    if ((rc = s->partial_delta.reserve((size_t)parts * n * s->row_bytes))) return rc;
    mg::PointParams p{};
    p.rows = s->d_cdiffs;
    p.row_idx = s->d_cidx;
    p.n_entries = c_n;
    p.chunks = s->chunks;
    p.keys = s->d_keys;
    p.key_idx = d_list;
    p.n_keys = n;
    p.partial = (uint4*)s->partial_delta.p;
    MG_HIP(mg::launch_point(p, st));
  }
  MG_HIP(mg::launch_reduce((const uint4*)s->partial_delta.p, parts, n, s->chunks, s->d_acc, d_list,
                           false, st));
  MG_HIP(hipStreamSynchronize(st));
  return MG_OK;
}

// One sweep step over segment s->step for every active key; returns the
// slots that have now seen every segment. Caller holds dev->mu.
int sweep_step(mg_server* s, std::vector<uint32_t>* finished) {
  hipStream_t st = s->dev->stream;
  const uint64_t row_begin = (uint64_t)s->step * s->seg_rows;
  const uint64_t row_end = std::min(s->n_rows, row_begin + s->seg_rows);
  uint32_t* h_list = (uint32_t*)s->h_idx.p;
  uint32_t off = 0;
  for (int b = 0; b < 2; b++) {
    const uint32_t start = off;
    for (uint32_t i = 0; i < s->cap_keys; i++)
      if (s->slots[i].req && s->slots[i].base == b) h_list[off++] = i;
    const uint32_t n = off - start;
    if (!n || row_end <= row_begin) continue;
    // The scan's own choice for this row width and key count (lane-DFS up to 128-byte rows,
    // then the stream or split kernel). Segments are multiples of the 2^(kl + lb)-row tile of
    // the loop's lane depth; lane-DFS rows let the plan pick the depth for this step's key
    // count among those whose tiles divide that (every smaller one does).
    mg_scan_opts o{};
    o.kernel = MG_KERNEL_AUTO;
    o.lane_depth = s->chunks <= 8 ? 0 : s->kl;
    o.row_begin = row_begin;
    o.row_end = row_end;
    mg::Plan pl;
    int rc = mg::make_plan(s->dev, s->base_rows[b], s->n_rows, s->row_bytes, n, s->d, &o, &pl);
    if (rc) return rc;
    if ((rc = s->partial.reserve((size_t)pl.parts * n * s->row_bytes))) return rc;
    if (pl.gslots && (rc = s->gstack.reserve(mg::lane_gstack_bytes(pl, n)))) return rc;
    MG_HIP(hipMemcpyAsync(s->d_idx + start, h_list + start, n * 4, hipMemcpyHostToDevice, st));
    mg::ScanBuffers bufs;
    bufs.keys = s->d_keys;
    bufs.key_idx = s->d_idx + start;
    bufs.partial = (uint4*)s->partial.p;
    bufs.gstack = (uint4*)s->gstack.p;
    bufs.out = s->d_acc;
    bufs.out_idx = s->d_idx + start;
    bufs.accumulate = true;
    if (s->use_graph) {
      // The plan with the buffers it reads and writes is the graph's signature.
      pl.ws_keys = bufs.keys;
      pl.ws_out = bufs.out;
      pl.ws_partial = bufs.partial;
      pl.ws_gstack = bufs.gstack;
      pl.ws_key_idx = bufs.key_idx;
      mg_server::StepGraph& g = s->graphs[(size_t)s->step * 2 + b];
      if (!(g.valid && g.plan == pl)) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = mg_server::StepGraph{};
        MG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = mg::enqueue_scan(s->dev, mg::control_exec(s->dev), pl, bufs, nullptr, nullptr);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc || e != hipSuccess) {
          if (graph) (void)hipGraphDestroy(graph);
          return rc ? rc : mg::hip_fail(e, "hipStreamEndCapture(sweep step)");
        }
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
          g.exec = nullptr;
          return mg::hip_fail(e, "hipGraphInstantiate(sweep step)");
        }
        g.plan = pl;
        g.valid = true;
      }
      MG_HIP(hipGraphLaunch(g.exec, st));
      s->graph_launches++;
    } else if ((rc = mg::enqueue_scan(s->dev, mg::control_exec(s->dev), pl, bufs, nullptr, nullptr))) {
      return rc;
    }
    // The slot list region is reused by the next base's launch only after
    // this one's kernels have read it.
    MG_HIP(hipStreamSynchronize(st));
    if ((rc = mg::check_fault(mg::control_exec(s->dev), pl))) return rc;
  }
  for (uint32_t i = 0; i < s->cap_keys; i++) {
    Slot& sl = s->slots[i];
    if (!sl.req) continue;
    if (++sl.steps_done == s->segments) finished->push_back(i);
  }
  s->step = (s->step + 1) % s->segments;
  if (!finished->empty()) {
    uint8_t* h = (uint8_t*)s->h_acc.p;
    for (size_t j = 0; j < finished->size(); j++)
      MG_HIP(hipMemcpyAsync(h + j * s->row_bytes, (const uint8_t*)s->d_acc +
                                                      (size_t)(*finished)[j] * s->row_bytes,
                            s->row_bytes, hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
  }
  return MG_OK;
}

void fail_all(mg_server* s, int code, const std::string& msg) {
  std::lock_guard<std::mutex> lk(s->mu);
  s->fatal = code;
  s->fatal_msg = msg;
  for (auto& kv : s->requests) {
    Request* r = kv.second.get();
    if (!r->done) {
      r->done = true;
      r->status = code;
      r->error = msg;
    }
  }
  // A queued snapshot's record lives on its caller's stack: complete it and
  // drop the event, so nothing touches the record once the caller returns.
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    if (it->snap) {
      it->snap->done = true;
      it->snap->status = code;
      it->snap->error = msg;
      it = s->queue.erase(it);
    } else {
      ++it;
    }
  }
  s->cv_done.notify_all();
}

// Rows per sweep step: whole tiles of the fused kernel (2^(kl + lb) rows),
// `want` steps per pass, or one step per 2^25 rows (at most 64) by default.
// Measured at 2^27 x 32 B under Poisson load, 4 steps per pass gave p50/p99
// of 416/1340 ms against 504/2003 ms at 32 steps: larger steps spend less on
// launches and host round trips, and admission waits at most one step.
void plan_segments(mg_server* s) {
  const uint32_t lb = std::min<uint32_t>(6, s->d - s->kl);
  const uint64_t tile = 1ull << (s->kl + lb);
  const uint32_t want =
      s->seg_want ? s->seg_want
                  : (uint32_t)std::min<uint64_t>(64, std::max<uint64_t>(1, s->n_rows >> 25));
  uint64_t seg = (s->n_rows + want - 1) / want;
  seg = std::max<uint64_t>(tile, (seg + tile - 1) / tile * tile);
  s->seg_rows = seg;
  s->segments = (uint32_t)((s->n_rows + seg - 1) / seg);
  drop_graphs(s);
  s->graphs.resize((size_t)s->segments * 2);
}

// Serves `m` from now on. Caller holds dev->mu; no key is in flight.
void apply_snapshot(mg_server* s, const Snapshot& sn) {
  const size_t bytes = (size_t)sn.m->n_rows * s->row_bytes;
  if (s->double_buffer && sn.m->n_rows != s->n_rows) {
    // A spare of the new size; without the HBM for it the loop goes on
    // single-buffered (merges then wait for in-flight keys).
    uint4* spare = nullptr;
    if (hipMalloc((void**)&spare, bytes) != hipSuccess) {
      spare = nullptr;
      (void)hipGetLastError();
    }
    (void)hipFree(s->base_rows[1]);
    s->base_rows[1] = spare;
    if (!spare) s->double_buffer = false;
  }
  s->m->users--;
  s->m = sn.m;
  s->base_rows[0] = sn.m->rows;
  s->cur_base = 0;
  s->n_rows = sn.m->n_rows;
  plan_segments(s);
  s->step = 0;
  s->c_idx.clear();
  s->c_diffs.clear();
  s->rebase_wanted = false;
}

// Frees the slots of requests abandoned in flight (their answers would be
// dropped anyway). Caller holds mu; the loop thread owns the slots.
void release_cancelled(mg_server* s) {
  for (uint32_t i = 0; i < s->cap_keys; i++) {
    Slot& sl = s->slots[i];
    if (!sl.req || !sl.req->cancel) continue;
    Request* r = sl.req;
    sl = Slot{};
    s->free_slots.push_back(i);
    s->active--;
    if (--r->keys_left == 0) {
      s->inflight.erase(std::find(s->inflight.begin(), s->inflight.end(), r));
      s->requests.erase(r->ticket);  // `r` is gone from here on
    }
  }
}

// Steps timed before the fitted model replaces the prior.
constexpr double kFitMinSteps = 4;

// Starts the step-time model afresh for the current segments, from a prior:
// per key, a segment's GGM nodes and folds (700 + W/4 int32 ops per row) at
// half the 78.6 TOP/s VALU peak; per step, the segment's bytes at 6.5 TB/s
// plus 0.1 ms of launches and host round trips. (2^31 x 32 B in 64 steps:
// 0.60 ms per key against 0.55 ms measured.) Caller holds mu, or owns `s`.
void reset_model(mg_server* s) {
  s->fit_w = s->fit_n = s->fit_t = s->fit_nn = s->fit_nt = 0;
  const double rows = (double)s->seg_rows;
  s->prior_key_ms = rows * (700.0 + s->row_bytes / 4.0) / 39.3e12 * 1e3;
  s->prior_fixed_ms = rows * s->row_bytes / 6.5e12 * 1e3 + 0.1;
  s->key_step_ms = s->prior_key_ms;
  s->step_fixed_ms = s->prior_fixed_ms;
}

// Refits step time = fixed + per-key x keys after a step of n keys that took t ms.
void fit_step(mg_server* s, double n, double t) {
  const double keep = 0.97;  // about the last 30 steps weigh in
  s->fit_w = keep * s->fit_w + 1;
  s->fit_n = keep * s->fit_n + n;
  s->fit_t = keep * s->fit_t + t;
  s->fit_nn = keep * s->fit_nn + n * n;
  s->fit_nt = keep * s->fit_nt + n * t;
  const double den = s->fit_w * s->fit_nn - s->fit_n * s->fit_n;
  double c = 0, a = 0;
  if (den > 1e-6 * s->fit_w * s->fit_nn) {
    c = (s->fit_w * s->fit_nt - s->fit_n * s->fit_t) / den;
    a = (s->fit_t - c * s->fit_n) / s->fit_w;
  }
  if (c <= 0 || a < 0) {  // too little spread in the key counts yet: all per key
    c = s->fit_t / s->fit_n;
    a = 0;
  }
  if (s->fit_w < kFitMinSteps) return;  // the prior stays in use
  s->key_step_ms = c;
  s->step_fixed_ms = a;
}

// Whether a request of n_keys submitted now, due at now + deadline_ms, can be
// admitted: with it, it and every request queued or in flight that would
// otherwise finish in time still finish in time. Requests queued are taken
// as admitted now (FIFO, ahead of it); each step costs fixed + per-key x the
// keys it serves. Caller holds mu.
bool admissible(mg_server* s, uint32_t n_keys, double now, double* due_out) {
  struct Job {
    uint32_t steps_left, keys;
    double due;
  };
  std::vector<Job> jobs;
  jobs.reserve(s->inflight.size() + s->queue.size() + 1);
  for (const Request* r : s->inflight)
    if (!r->cancel) jobs.push_back({s->segments - std::min(r->steps_done, s->segments), r->n_keys,
                                    r->due});
  for (const Event& ev : s->queue)
    if (ev.req) jobs.push_back({s->segments, ev.req->n_keys, ev.req->due});
  // Keys served by step j (1-based) of the schedule, without the new request.
  std::vector<double> keys_at(s->segments + 2, 0.0);
  for (const Job& j : jobs) keys_at[j.steps_left] += j.keys;
  for (int j = (int)s->segments - 1; j >= 1; --j) keys_at[j] += keys_at[j + 1];
  // Finish times without (t0) and with (t1) the new request's keys in steps 1..S.
  std::vector<double> t0(s->segments + 1, now), t1(s->segments + 1, now);
  for (uint32_t j = 1; j <= s->segments; ++j) {
    t0[j] = t0[j - 1] + s->step_fixed_ms + s->key_step_ms * keys_at[j];
    t1[j] = t1[j - 1] + s->step_fixed_ms + s->key_step_ms * (keys_at[j] + n_keys);
  }
  const double due = now + s->deadline_ms;
  *due_out = t1[s->segments];
  if (t1[s->segments] > due) return false;
  for (const Job& j : jobs)
    if (j.due > 0 && t1[j.steps_left] > j.due && t0[j.steps_left] <= j.due) return false;
  return true;
}

void server_loop(mg_server* s) {
  (void)hipSetDevice(s->dev->ordinal);
  // One turn's work, in submission order: groups of requests to admit and
  // updates between them.
  struct Item {
    std::vector<Request*> admit;
    std::unique_ptr<Update> upd;
  };
  for (;;) {
    std::vector<Item> items;
    Snapshot* snap = nullptr;  // taken only when no key is in flight
    bool snap_applied = false;  // once set, `snap` may already be gone
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        return s->stop || !s->queue.empty() || s->active > 0 || s->rebase_wanted ||
               s->inject != MG_OK;
      });
      if (s->stop) break;
      if (s->inject != MG_OK) {
        const int code = s->inject;
        s->inject = MG_OK;
        lk.unlock();
        fail_all(s, code, "injected failure");
        break;
      }
      release_cancelled(s);
      // Events up to the first snapshot (which needs every key before it
      // done), in order: requests are admitted FIFO while their keys fit; once
      // one does not, later requests wait too, but updates behind it are still
      // taken: an update waits for no request, and the requests it overtakes
      // answer at its epoch. An update the delta has no room for (a rebase
      // makes room first) holds back everything behind it.
      // Single-buffered: a pending in-place merge stops admissions.
      bool blocked = s->rebase_wanted && !s->double_buffer;
      uint32_t free_n = (uint32_t)s->free_slots.size();
      uint64_t c_n = s->c_idx.size();
      for (auto it = s->queue.begin(); it != s->queue.end();) {
        if (it->snap) break;
        if (it->upd) {
          if (c_n + it->upd->idx.size() > s->c_cap) {
            s->rebase_wanted = true;
            break;
          }
          c_n += it->upd->idx.size();
          items.emplace_back();
          items.back().upd = std::move(it->upd);
          it = s->queue.erase(it);
          continue;
        }
        if (blocked || it->req->n_keys > free_n) {
          blocked = true;
          ++it;
          continue;
        }
        free_n -= it->req->n_keys;
        if (items.empty() || items.back().upd) items.emplace_back();
        items.back().admit.push_back(it->req);
        s->queued_keys -= it->req->n_keys;
        // Out of the queue means admitted for mg_server_abandon: between here and the
        // slots below `mu` is released, and an abandon that took the request for a queued
        // one would free it under the loop (it now cancels it, and the next step releases
        // its slots).
        it->req->admitted = true;
        it = s->queue.erase(it);
      }
      if (items.empty() && s->active == 0 && !s->queue.empty() && s->queue.front().snap) {
        snap = s->queue.front().snap;
        s->queue.pop_front();
      }
    }
    const double t0 = mg::now_ms();
    std::vector<uint32_t> finished;
    int rc = MG_OK;
    uint32_t n_stepped = 0;
    {
      std::lock_guard<std::mutex> dlk(s->dev->mu);
      rc = mg::set_device(s->dev);
      if (!rc && !snap) rc = try_rebase(s);
      for (Item& item : items) {
        if (rc) break;
        if (item.upd) {
          rc = apply_update(s, *item.upd);
          if (!rc) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->epoch = item.upd->epoch;
          }
          continue;
        }
        const double now = mg::now_ms();
        uint32_t n_new = 0;
        {
          std::lock_guard<std::mutex> lk(s->mu);
          for (Request* r : item.admit) {
            r->t_admit = now;
            r->epoch = s->epoch;
            r->admitted = true;
            r->keys_left = r->n_keys;
            n_new += r->n_keys;
            s->inflight.push_back(r);
          }
        }
        rc = admit_keys(s, item.admit);
        if (!rc) s->active += n_new;
      }
      if (!rc && snap) {
        apply_snapshot(s, *snap);
        std::lock_guard<std::mutex> lk(s->mu);
        s->epoch = snap->epoch;
        s->stats.segments = s->segments;
        s->stats.snapshots++;
        s->stats.n_rows = s->n_rows;
        reset_model(s);
        snap->done = true;
        snap_applied = true;
        s->cv_done.notify_all();
      }
      n_stepped = s->active;
      if (!rc && s->active) rc = sweep_step(s, &finished);
    }
    if (snap && !snap_applied) {  // the loop failed before the swap
      std::lock_guard<std::mutex> lk(s->mu);
      snap->done = true;
      snap->status = rc ? rc : MG_ERR_STATE;
      snap->error = mg::last_error();
      s->cv_done.notify_all();
    }
    if (rc) {
      fail_all(s, rc, mg::last_error());
      break;
    }
    const double t1 = mg::now_ms();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stats.busy_ms += t1 - t0;
      if (n_stepped) {
        s->stats.sweep_steps++;
        for (Request* r : s->inflight) r->steps_done++;
        // The admission model's step time (steps that also admitted, applied
        // updates or merged count in full: such work delays requests too).
        fit_step(s, n_stepped, t1 - t0);
      }
      const uint8_t* h = (const uint8_t*)s->h_acc.p;
      for (size_t j = 0; j < finished.size(); j++) {
        Slot& sl = s->slots[finished[j]];
        Request* r = sl.req;
        memcpy(r->answers.data() + (size_t)sl.key * s->row_bytes, h + j * s->row_bytes,
               s->row_bytes);
        r->keys_done++;
        if (--r->keys_left == 0)
          s->inflight.erase(std::find(s->inflight.begin(), s->inflight.end(), r));
        if (r->keys_done == r->n_keys) {
          r->done = true;
          r->t_done = t1;
          s->stats.completed++;
          s->stats.keys_completed += r->n_keys;
          if (r->abandoned) s->requests.erase(r->ticket);  // `r` is gone from here on
        } else if (r->cancel && r->keys_left == 0) {
          s->requests.erase(r->ticket);
        }
        sl = Slot{};
        s->free_slots.push_back(finished[j]);
        s->active--;
      }
      s->stats.active_keys = s->active;
      s->stats.delta_entries = s->c_idx.size();
      s->stats.rebases = s->rebases;
      s->stats.graph_launches = s->graph_launches;
      if (!finished.empty()) s->cv_done.notify_all();
    }
  }
}

}  // namespace

namespace mg {
uint32_t server_row_bytes(const mg_server* s) { return s->row_bytes; }
}  // namespace mg

extern "C" {

int mg_server_create(mg_device* dev, mg_matrix* m, const mg_server_opts* opts, mg_server** out) {
  if (!dev || !m || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
  if (m->dev != dev) return fail(MG_ERR_INVALID_ARG, "matrix belongs to another device");
  *out = nullptr;
  mg_server_opts def{};
  const mg_server_opts* o = opts ? opts : &def;
  if (m->n_rows == 0) return fail(MG_ERR_ROWS, "empty matrix");
  uint32_t d = o->domain_bits;
  if (!d) d = std::max<uint32_t>(1, 64 - __builtin_clzll(m->n_rows - 1 ? m->n_rows - 1 : 1));
  if (d > (uint32_t)mg::kMaxDomainBits) return fail(MG_ERR_DOMAIN, "domain bits %u > 32", d);
  if (d < 64 && m->n_rows > (1ull << d))
    return fail(MG_ERR_ROWS, "%llu rows exceed the 2^%u-leaf domain",
                (unsigned long long)m->n_rows, d);
  if (m->row_bytes > 4096) return fail(MG_ERR_ROWS, "server rows are at most 4096 bytes");

  std::unique_ptr<mg_server> s(new mg_server());
  s->dev = dev;
  s->m = m;
  s->row_bytes = m->row_bytes;
  s->chunks = m->row_bytes / 16;
  s->d = d;
  s->n_rows = m->n_rows;
  s->double_buffer = o->double_buffer != 0;
  s->deadline_ms = o->deadline_ms;
  s->use_graph = o->use_graph != 0;
  s->kl = std::min<uint32_t>(o->lane_depth ? o->lane_depth : (s->chunks > 8 ? 7 : 10), d);
  s->seg_want = o->segments;
  s->queued_n_rows = s->n_rows;
  plan_segments(s.get());
  reset_model(s.get());
  s->c_cap = o->max_delta_entries ? o->max_delta_entries : (1ull << 20);

  std::lock_guard<std::mutex> lk(dev->mu);
  int rc = mg::set_device(dev);
  if (rc) return rc;
  size_t free_b = 0, total_b = 0;
  MG_HIP(hipMemGetInfo(&free_b, &total_b));
  const size_t matrix_bytes = (size_t)m->n_rows * m->row_bytes;
  const size_t delta_bytes = s->c_cap * (8 + (size_t)m->row_bytes);
  size_t need = delta_bytes + (s->double_buffer ? matrix_bytes : 0) + o->hbm_reserve_bytes;
  if (need > free_b)
    return fail(MG_ERR_OUT_OF_MEMORY, "server needs %zu B besides the matrix, %zu B free", need,
                free_b);
  // Per key: its record, accumulator, two slot-list entries and a share of
  // the partial buffers; 64 MiB is kept for workspace growth.
  const size_t per_key = sizeof(DevKey) * 2 + 3 * (size_t)m->row_bytes + 8;
  size_t budget = free_b - need;
  budget = budget > (64u << 20) ? budget - (64u << 20) : 0;
  uint64_t cap = std::min<uint64_t>(budget / per_key, 32768);
  if (o->max_active_keys) cap = std::min<uint64_t>(cap, o->max_active_keys);
  if (cap == 0) return fail(MG_ERR_OUT_OF_MEMORY, "no HBM left for key slots");
  s->cap_keys = (uint32_t)cap;

  hipError_t e = hipMalloc((void**)&s->d_cidx, s->c_cap * 8);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_cdiffs, s->c_cap * m->row_bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_keys, cap * sizeof(DevKey));
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_acc, cap * m->row_bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_idx, cap * 2 * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_stage_keys, cap * sizeof(DevKey));
  if (e == hipSuccess && s->double_buffer) e = hipMalloc((void**)&s->base_rows[1], matrix_bytes);
  if (e != hipSuccess) {
    free_server_memory(s.get());
    return mg::hip_fail(e, "hipMalloc(server)");
  }
  if ((rc = s->h_idx.reserve(cap * 2 * 4)) || (rc = s->h_keys.reserve(cap * sizeof(DevKey))) ||
      (rc = s->h_acc.reserve(cap * m->row_bytes))) {
    free_server_memory(s.get());
    return rc;
  }
  s->base_rows[0] = m->rows;
  s->slots.assign(cap, Slot{});
  s->free_slots.reserve(cap);
  for (uint32_t i = (uint32_t)cap; i-- > 0;) s->free_slots.push_back(i);
  s->stats.max_active_keys = (uint32_t)cap;
  s->stats.segments = s->segments;
  s->stats.n_rows = s->n_rows;
  m->users++;
  dev->children++;
  s->thread = std::thread(server_loop, s.get());
  *out = s.release();
  return MG_OK;
}

int mg_server_submit(mg_server* s, const uint8_t* keys, const size_t* key_lens, uint32_t n_keys,
                     uint64_t* ticket) {
  if (!s || !ticket || (n_keys && (!keys || !key_lens)))
    return fail(MG_ERR_INVALID_ARG, "null argument");
  if (n_keys == 0) return fail(MG_ERR_INVALID_ARG, "a request needs at least one key");
  if (n_keys > s->cap_keys)
    return fail(MG_ERR_BUSY, "request of %u keys exceeds the %u key slots", n_keys, s->cap_keys);
  std::unique_ptr<Request> r(new Request());
  r->keys.resize(n_keys);
  uint32_t d = 0;
  int rc = mg::parse_keys(keys, key_lens, n_keys, r->keys.data(), &d);
  if (rc) return rc;
  if (d != s->d) return fail(MG_ERR_DOMAIN, "keys have domain %u, the server serves %u", d, s->d);
  r->n_keys = n_keys;
  r->answers.assign((size_t)n_keys * s->row_bytes, 0);
  r->t_submit = mg::now_ms();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->fatal) return fail(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
  if (s->deadline_ms) {
    r->due = r->t_submit + s->deadline_ms;
    double finish = 0;
    if (!admissible(s, n_keys, r->t_submit, &finish)) {
      s->stats.rejected++;
      return fail(MG_ERR_BUSY, "request of %u keys would finish in about %.0f ms, or delay "
                               "requests admitted before it, past the %u ms deadline",
                  n_keys, finish - r->t_submit, s->deadline_ms);
    }
  }
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  Event ev;
  ev.req = r.get();
  s->queue.push_back(std::move(ev));
  s->queued_keys += n_keys;
  s->requests[r->ticket] = std::move(r);
  s->stats.submitted++;
  s->cv_work.notify_one();
  return MG_OK;
}

int mg_server_wait(mg_server* s, uint64_t ticket, uint8_t* out, uint32_t timeout_ms,
                   mg_request_info* info) {
  if (!s) return fail(MG_ERR_INVALID_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) return fail(MG_ERR_INVALID_ARG, "unknown ticket %llu",
                                           (unsigned long long)ticket);
  Request* r = it->second.get();
  if (r->abandoned) return fail(MG_ERR_INVALID_ARG, "ticket %llu was abandoned",
                                (unsigned long long)ticket);
  if (r->waiting)
    return fail(MG_ERR_STATE, "ticket %llu already has a thread waiting for it",
                (unsigned long long)ticket);
  auto ready = [&] { return r->done; };
  s->waiters++;
  r->waiting++;  // abandon() refuses the ticket meanwhile, so `r` stays valid
  bool ok = true;
  if (timeout_ms)
    ok = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
  else
    s->cv_done.wait(lk, ready);
  r->waiting--;
  if (--s->waiters == 0) s->cv_gone.notify_all();
  if (!ok)
    return fail(MG_ERR_BUSY, "request %llu not done after %u ms", (unsigned long long)ticket,
                timeout_ms);
  if (r->status != MG_OK) {
    const int code = r->status;
    const std::string msg = r->error;
    s->requests.erase(it);
    return fail(code, "request failed: %s", msg.c_str());
  }
  if (out) memcpy(out, r->answers.data(), r->answers.size());
  if (info) {
    info->queue_ms = r->t_admit - r->t_submit;
    info->service_ms = r->t_done - r->t_admit;
    info->latency_ms = r->t_done - r->t_submit;
    info->epoch = r->epoch;
    info->n_keys = r->n_keys;
  }
  s->requests.erase(it);
  return MG_OK;
}

int mg_server_abandon(mg_server* s, uint64_t ticket) {
  if (!s) return fail(MG_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end() || it->second->abandoned)
    return fail(MG_ERR_INVALID_ARG, "unknown ticket %llu", (unsigned long long)ticket);
  Request* r = it->second.get();
  if (r->waiting)
    return fail(MG_ERR_STATE, "ticket %llu has a thread waiting for it", (unsigned long long)ticket);
  if (r->done) {
    s->requests.erase(it);
    return MG_OK;
  }
  s->stats.abandoned++;
  if (!r->admitted) {  // still queued: it never runs
    for (auto q = s->queue.begin(); q != s->queue.end(); ++q)
      if (q->req == r) {
        s->queue.erase(q);
        break;
      }
    s->queued_keys -= r->n_keys;
    s->requests.erase(it);
    return MG_OK;
  }
  // In flight: its keys leave the sweep at the loop's next step.
  r->abandoned = true;
  r->cancel = true;
  s->cv_work.notify_one();
  return MG_OK;
}

int mg_server_update(mg_server* s, const uint64_t* row_idx, const uint8_t* diffs,
                     uint64_t n_entries, uint64_t epoch) {
  if (!s || (n_entries && (!row_idx || !diffs))) return fail(MG_ERR_INVALID_ARG, "null argument");
  if (n_entries > s->c_cap)
    return fail(MG_ERR_INVALID_ARG, "%llu update rows exceed the delta capacity %llu",
                (unsigned long long)n_entries, (unsigned long long)s->c_cap);
  std::unique_ptr<Update> u(new Update());
  u->idx.assign(row_idx, row_idx + n_entries);
  u->diffs.assign(diffs, diffs + n_entries * s->row_bytes);
  u->epoch = epoch;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->fatal) return fail(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
  // Rows as of the queue's end: an update applies to the matrix of the last
  // snapshot submitted before it.
  for (uint64_t i : u->idx)
    if (i >= s->queued_n_rows)
      return fail(MG_ERR_ROWS, "update row %llu outside the %llu-row matrix",
                  (unsigned long long)i, (unsigned long long)s->queued_n_rows);
  Event ev;
  ev.upd = std::move(u);
  s->queue.push_back(std::move(ev));
  s->cv_work.notify_one();
  return MG_OK;
}

int mg_server_snapshot(mg_server* s, mg_matrix* m, uint64_t epoch) {
  if (!s || !m) return fail(MG_ERR_INVALID_ARG, "null argument");
  if (m->dev != s->dev) return fail(MG_ERR_INVALID_ARG, "snapshot matrix is on another device");
  if (m->row_bytes != s->row_bytes)
    return fail(MG_ERR_ROWS, "snapshot rows are %u bytes, the loop serves %u", m->row_bytes,
                s->row_bytes);
  if (m->n_rows == 0 || (s->d < 64 && m->n_rows > (1ull << s->d)))
    return fail(MG_ERR_ROWS, "snapshot of %llu rows does not fit the 2^%u-leaf domain",
                (unsigned long long)m->n_rows, s->d);
  Snapshot sn;
  sn.m = m;
  sn.epoch = epoch;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->fatal) return fail(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
  m->users++;  // held from the queue on; the swap releases the old matrix
  Event ev;
  ev.snap = &sn;
  s->queue.push_back(std::move(ev));
  s->queued_n_rows = m->n_rows;
  s->cv_work.notify_one();
  s->waiters++;
  s->cv_done.wait(lk, [&] { return sn.done; });
  if (--s->waiters == 0) s->cv_gone.notify_all();
  if (sn.status != MG_OK) {
    m->users--;
    return fail(sn.status, "snapshot failed: %s", sn.error.c_str());
  }
  return MG_OK;
}

int mg_server_stats_get(mg_server* s, mg_server_stats* out) {
  if (!s || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->epoch = s->epoch;
  out->key_step_ms = s->key_step_ms;
  out->step_fixed_ms = s->step_fixed_ms;
  uint32_t q = 0;
  for (const Event& ev : s->queue)
    if (ev.req) q++;
  out->queued_requests = q;
  return MG_OK;
}

int mg_server_inject_failure(mg_server* s, int code) {
  if (!s || code == MG_OK) return fail(MG_ERR_INVALID_ARG, "null server or MG_OK code");
  std::lock_guard<std::mutex> lk(s->mu);
  s->inject = code;
  s->cv_work.notify_one();
  return MG_OK;
}

int mg_server_destroy(mg_server* s) {
  if (!s) return MG_OK;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& kv : s->requests)
      if (!kv.second->done) {
        kv.second->done = true;
        kv.second->status = MG_ERR_STATE;
        kv.second->error = "server destroyed";
      }
    // Queued snapshots: complete them and drop their events (the records live
    // on their callers' stacks).
    for (auto it = s->queue.begin(); it != s->queue.end();) {
      if (it->snap) {
        it->snap->done = true;
        it->snap->status = MG_ERR_STATE;
        it->snap->error = "server destroyed";
        it = s->queue.erase(it);
      } else {
        ++it;
      }
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    // Callers still inside mg_server_wait / mg_server_snapshot read the
    // server's state when they wake: free it only after they have left.
    s->cv_gone.wait(lk, [&] { return s->waiters == 0; });
  }
  if (s->thread.joinable()) s->thread.join();
  {
    std::lock_guard<std::mutex> dlk(s->dev->mu);
    std::unique_lock<mg::ScanGate> wlk(s->dev->state);
    (void)mg::set_device(s->dev);
    (void)hipStreamSynchronize(s->dev->stream);
    // A double-buffered server may have left the newest rows in its spare copy.
    if (s->cur_base == 1)
      (void)hipMemcpy(s->m->rows, s->base_rows[1], (size_t)s->n_rows * s->row_bytes,
                      hipMemcpyDeviceToDevice);
    free_server_memory(s);
  }
  s->m->users--;
  s->dev->children--;
  delete s;
  return MG_OK;
}

}  // extern "C"
