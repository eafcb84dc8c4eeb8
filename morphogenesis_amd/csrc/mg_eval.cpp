// Key evaluation on the device: ChaChaKey::{eval, eval_subtree, full_eval}
// (crates/morphogen-gpu-dpf/src/dpf.rs:103-246) behind the C ABI, with the
// reference's argument rules. Outputs are 16-byte leaves (Seed128, four
// little-endian words) in index order, produced in chunks of at most 2^24
// leaves (256 MiB) through a device staging buffer.
#include <algorithm>
#include <cstring>

#include "mg_internal.hpp"

namespace {

constexpr uint64_t kChunkLeaves = 1ull << 24;
constexpr uint64_t kChunkPoints = 1ull << 22;

// Uploads one parsed key to a workspace the caller holds.
int stage_key(mg::Workspace* w, const uint8_t* key, size_t key_len, mg::DevKey** d_key,
              uint32_t* d) {
  size_t len = key_len;
  int rc = w->h_keys.reserve(sizeof(mg::DevKey));
  if (rc) return rc;
  mg::DevKey* k = (mg::DevKey*)w->h_keys.p;
  if ((rc = mg::parse_keys(key, &len, 1, k, d))) return rc;
  if ((rc = w->keys.reserve(sizeof(mg::DevKey)))) return rc;
  MG_HIP(hipMemcpyAsync(w->keys.p, k, sizeof *k, hipMemcpyHostToDevice, w->stream));
  *d_key = (mg::DevKey*)w->keys.p;
  return MG_OK;
}

// Leaves [start, start + size) of an aligned subtree (size a power of two,
// start a multiple of size, size <= 2^d) into host `out`, on a workspace the
// caller holds.
int eval_range_to_host(mg::Workspace* w, const mg::DevKey* d_key, uint32_t d, uint64_t start,
                       uint64_t size, uint8_t* out) {
  hipStream_t s = w->stream;
  const uint64_t chunk = std::min(size, kChunkLeaves);
  int rc = w->out.reserve(chunk * 16);
  if (rc) return rc;
  for (uint64_t off = 0; off < size; off += chunk) {
    // A chunk is itself an aligned subtree of 2^c leaves; lanes take 2^kl
    // leaves each, with at least 64 lanes where the chunk allows it.
    const int c = __builtin_ctzll(chunk);
    const int kl = std::min<int>(std::min<int>(10, (int)d), std::max(0, c - 6));
    MG_HIP(mg::launch_eval_range(d_key, start + off, chunk >> kl, kl, (uint4*)w->out.p, s));
    MG_HIP(hipMemcpyAsync(out + off * 16, w->out.p, chunk * 16, hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
  }
  return MG_OK;
}

}  // namespace

extern "C" {

int mg_key_eval(mg_device* dev, const uint8_t* key, size_t key_len, const uint64_t* indices,
                uint64_t n, uint8_t* out) {
  if (!dev || !key || ((!indices || !out) && n)) return mg::fail(MG_ERR_INVALID_ARG, "null argument");
  int rc = mg::set_device(dev);
  if (rc) return rc;
  mg::WorkspaceLease w(dev);
  if ((rc = w->ready())) return rc;
  mg::DevKey* d_key = nullptr;
  uint32_t d = 0;
  if ((rc = stage_key(w.get(), key, key_len, &d_key, &d))) return rc;
  hipStream_t s = w->stream;
  if (!n) {
    MG_HIP(hipStreamSynchronize(s));  // the key upload reads the staging
    return MG_OK;
  }
  const uint64_t chunk = std::min(n, kChunkPoints);
  if ((rc = w->partial.reserve(chunk * 8)) || (rc = w->out.reserve(chunk * 16))) return rc;
  for (uint64_t off = 0; off < n; off += chunk) {
    const uint64_t m = std::min(chunk, n - off);
    MG_HIP(hipMemcpyAsync(w->partial.p, indices + off, m * 8, hipMemcpyHostToDevice, s));
    MG_HIP(mg::launch_eval_points(d_key, (const uint64_t*)w->partial.p, m, (uint4*)w->out.p, s));
    MG_HIP(hipMemcpyAsync(out + off * 16, w->out.p, m * 16, hipMemcpyDeviceToHost, s));
    MG_HIP(hipStreamSynchronize(s));
  }
  return MG_OK;
}

int mg_key_eval_subtree(mg_device* dev, const uint8_t* key, size_t key_len, uint64_t start,
                        uint64_t size, uint8_t* out) {
  if (!dev || !key || (!out && size)) return mg::fail(MG_ERR_INVALID_ARG, "null argument");
  // dpf.rs:164-175: a power-of-two size and a start aligned to it.
  if (size == 0 || (size & (size - 1)))
    return mg::fail(MG_ERR_KEY_FORMAT, "Subtree size must be power of 2");
  if (start % size) return mg::fail(MG_ERR_KEY_FORMAT, "Start index must be aligned to subtree size");
  int rc = mg::set_device(dev);
  if (rc) return rc;
  mg::WorkspaceLease w(dev);
  if ((rc = w->ready())) return rc;
  mg::DevKey* d_key = nullptr;
  uint32_t d = 0;
  if ((rc = stage_key(w.get(), key, key_len, &d_key, &d))) return rc;
  if (d < 64 && size > (1ull << d)) {
    (void)hipStreamSynchronize(w->stream);  // the key upload reads the staging
    return mg::fail(MG_ERR_ROWS, "a subtree of %llu leaves exceeds the 2^%u-leaf domain",
                    (unsigned long long)size, d);
  }
  return eval_range_to_host(w.get(), d_key, d, start, size, out);
}

int mg_key_full_eval(mg_device* dev, const uint8_t* key, size_t key_len, uint8_t* out,
                     uint64_t out_len) {
  if (!dev || !key || !out) return mg::fail(MG_ERR_INVALID_ARG, "null argument");
  mg_key_info info;
  int rc = mg_key_parse(key, key_len, &info);
  if (rc) return rc;
  const uint64_t leaves = 1ull << info.domain_bits;
  // dpf.rs:147-158: the output holds exactly 2^d leaves (InvalidKeyLength).
  if (out_len != leaves * 16)
    return mg::fail(MG_ERR_INVALID_ARG, "full_eval needs %llu output bytes, got %llu",
                    (unsigned long long)(leaves * 16), (unsigned long long)out_len);
  return mg_key_eval_subtree(dev, key, key_len, 0, leaves, out);
}

}  // extern "C"
