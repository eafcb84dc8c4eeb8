// mg_scan_noalloc: the fused scan with every per-call device buffer carved
// out of one fixed device array, for callers that cannot allocate.
//
// The array below is a file-scope __device__ object, so each GPU the library
// runs on holds its own copy from code-object load until the process ends.
// Its name is only a host-side stand-in: host code uses the per-device address
// that mg_device_open resolves once (noalloc_setup) and keeps in
// mg_device::noalloc_scratch. Nothing in it has to start at zero: every byte a
// scan reads there (keys, per-wave partials, stack levels, answers) is written
// earlier in the same call.
#include <algorithm>
#include <cstring>

#include "mg_internal.hpp"

static __device__ __align__(256) uint4 scratch[MG_NOALLOC_SCRATCH_BYTES / 16];

namespace mg {

int noalloc_setup(mg_device* dev) {
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  if (!p) return fail(MG_ERR_DEVICE, "no device address for the no-alloc scratch array");
  dev->noalloc_scratch = p;
  dev->noalloc_keys.resize(MG_NOALLOC_MAX_KEYS);
  return MG_OK;
}

}  // namespace mg

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Byte offsets of one scan's buffers inside the scratch array.
struct Layout {
  size_t keys = 0, out = 0, partial = 0, partial_delta = 0, gstack = 0, end = 0;
};

Layout layout_for(const mg::Plan& pl, uint32_t n_keys, uint32_t row_bytes) {
  Layout l;
  size_t off = 0;
  l.keys = off;
  off = align_up(off + sizeof(mg::DevKey) * n_keys);
  l.out = off;
  off = align_up(off + (size_t)n_keys * row_bytes);
  l.partial = off;
  off = align_up(off + (size_t)std::max<uint32_t>(pl.parts, 1) * n_keys * row_bytes);
  l.partial_delta = off;
  if (pl.n_delta) off = align_up(off + (size_t)pl.delta_parts * n_keys * row_bytes);
  l.gstack = off;
  off = align_up(off + mg::lane_gstack_bytes(pl, n_keys));
  l.end = off;
  return l;
}

}  // namespace

extern "C" int mg_scan_noalloc(mg_device* dev, mg_matrix* m, const uint8_t* keys,
                               const size_t* key_lens, uint32_t n_keys, uint8_t* out,
                               const mg_scan_opts* opts, mg_timing* timing) {
  using mg::fail;
  if (!dev || !m) return fail(MG_ERR_INVALID_ARG, "null device or matrix");
  if (m->dev != dev) return fail(MG_ERR_INVALID_ARG, "matrix belongs to another device");
  if (n_keys == 0) return MG_OK;
  if (!keys || !key_lens || !out) return fail(MG_ERR_INVALID_ARG, "null keys or output");
  if (n_keys > MG_NOALLOC_MAX_KEYS)
    return fail(MG_ERR_INVALID_ARG, "%u keys exceed the no-alloc limit of %u", n_keys,
                MG_NOALLOC_MAX_KEYS);
  if (opts && opts->use_graph)
    return fail(MG_ERR_INVALID_ARG, "use_graph is not available without allocation");
  std::lock_guard<std::mutex> lk(dev->mu);
  int rc = mg::set_device(dev);
  if (rc) return rc;
  if (!dev->noalloc_scratch || dev->noalloc_keys.size() < MG_NOALLOC_MAX_KEYS)
    return fail(MG_ERR_STATE, "the no-alloc scratch array was not set up for this device");
  const double t0 = mg::now_ms();
  uint32_t d = 0;
  if ((rc = mg::parse_keys(keys, key_lens, n_keys, dev->noalloc_keys.data(), &d))) return rc;

  mg::Plan pl;
  if ((rc = mg::make_plan(dev, m->rows, m->n_rows, m->row_bytes, n_keys, d, opts, &pl))) return rc;
  if ((rc = mg::attach_delta(dev, m, opts, &pl))) return rc;
  Layout l = layout_for(pl, n_keys, m->row_bytes);
  // Lane-DFS stack levels in global memory grow with the waves; when the
  // planner's choice does not fit and the caller left the levels open, keep
  // every level in LDS instead (the same answers, one plan further down the
  // planner's list).
  if (l.end > MG_NOALLOC_SCRATCH_BYTES && pl.gslots &&
      (!opts || opts->stack_global == 0)) {
    mg_scan_opts o2 = opts ? *opts : mg_scan_opts{};
    o2.stack_global = 1;
    pl = mg::Plan();
    if ((rc = mg::make_plan(dev, m->rows, m->n_rows, m->row_bytes, n_keys, d, &o2, &pl)))
      return rc;
    if ((rc = mg::attach_delta(dev, m, &o2, &pl))) return rc;
    l = layout_for(pl, n_keys, m->row_bytes);
  }
  if (l.end > MG_NOALLOC_SCRATCH_BYTES)
    return fail(MG_ERR_INVALID_ARG,
                "this scan needs %zu B of device scratch; the no-alloc path has %u B", l.end,
                MG_NOALLOC_SCRATCH_BYTES);

  uint8_t* base = (uint8_t*)dev->noalloc_scratch;
  mg::ScanBuffers b;
  b.keys = (const mg::DevKey*)(base + l.keys);
  b.out = (uint4*)(base + l.out);
  b.partial = (uint4*)(base + l.partial);
  b.partial_delta = pl.n_delta ? (uint4*)(base + l.partial_delta) : nullptr;
  b.gstack = pl.gslots ? (uint4*)(base + l.gstack) : nullptr;
  const size_t out_bytes = (size_t)n_keys * m->row_bytes;

  hipStream_t s = dev->stream;
  MG_HIP(hipEventRecord(dev->ev[0], s));
  MG_HIP(hipMemcpyAsync(base + l.keys, dev->noalloc_keys.data(), sizeof(mg::DevKey) * n_keys,
                        hipMemcpyHostToDevice, s));
  if ((rc = mg::enqueue_scan(dev, mg::control_exec(dev), pl, b, dev->ev[1], dev->ev[2])))
    return rc;
  MG_HIP(hipEventRecord(dev->ev[3], s));
  MG_HIP(hipMemcpyAsync(out, base + l.out, out_bytes, hipMemcpyDeviceToHost, s));
  MG_HIP(hipEventRecord(dev->ev[4], s));
  MG_HIP(hipStreamSynchronize(s));
  if ((rc = mg::check_fault(mg::control_exec(dev), pl))) return rc;
  if (timing) return mg::fill_timing(dev->ev, pl, t0, timing);
  return MG_OK;
}
