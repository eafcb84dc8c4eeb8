// Raw row file -> device matrix.
//
// Replaces the reference's matrix ingestion: load_matrix_from_file
// (crates/morphogen-server/src/bin/bench_page_gpu_batch.rs:326-379) followed
// by GpuPageMatrix::from_chunked_matrix (crates/morphogen-gpu-dpf/src/
// storage.rs:46-86), which read the file into host chunks, zero-pad the tail
// and upload chunk by chunk; and the snapshot decode rule
// (crates/morphogen-server/src/network/snapshot_config.rs:144-159), which
// refuses an empty file or one that is not a whole number of rows.
//
// Here the file streams straight into HBM: a reader fills one of two pinned
// staging buffers while the other one's copy runs on a stream of its own, so
// disk (or page cache) reads overlap the PCIe copies and the device's
// serving stream is never blocked. The matrix is zeroed at creation, so a
// ragged tail is zero-padded by copying only the file's bytes.
#include <errno.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>

#include "mg_internal.hpp"

namespace {

constexpr size_t kStageBytes = 64u << 20;

// Reads exactly n bytes at off (short reads and EINTR retried).
bool read_at(int fd, uint8_t* dst, size_t n, uint64_t off) {
  while (n) {
    const ssize_t r = pread(fd, dst, n, (off_t)off);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    dst += r;
    n -= (size_t)r;
    off += (uint64_t)r;
  }
  return true;
}

struct Staging {
  hipStream_t stream = nullptr;
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  ~Staging() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (int i = 0; i < 2; i++) {
      if (ev[i]) (void)hipEventDestroy(ev[i]);
      if (buf[i]) (void)hipHostFree(buf[i]);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

int stream_file(mg_matrix* m, int fd, uint64_t file_bytes) {
  Staging st;
  MG_HIP(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
  const size_t stage = (size_t)std::min<uint64_t>(kStageBytes, file_bytes);
  for (int i = 0; i < 2; i++) {
    MG_HIP(hipHostMalloc(&st.buf[i], stage, hipHostMallocDefault));
    MG_HIP(hipEventCreateWithFlags(&st.ev[i], hipEventDisableTiming));
  }
  uint8_t* dst = (uint8_t*)m->rows;
  int b = 0;
  for (uint64_t off = 0; off < file_bytes; off += stage, b ^= 1) {
    const size_t n = (size_t)std::min<uint64_t>(stage, file_bytes - off);
    if (st.used[b]) MG_HIP(hipEventSynchronize(st.ev[b]));  // its previous copy is done
    if (!read_at(fd, (uint8_t*)st.buf[b], n, off))
      return mg::fail(MG_ERR_ROWS, "reading %zu bytes at offset %llu failed", n,
                      (unsigned long long)off);
    MG_HIP(hipMemcpyAsync(dst + off, st.buf[b], n, hipMemcpyHostToDevice, st.stream));
    MG_HIP(hipEventRecord(st.ev[b], st.stream));
    st.used[b] = true;
  }
  MG_HIP(hipStreamSynchronize(st.stream));
  return MG_OK;
}

}  // namespace

extern "C" int mg_matrix_load_file(mg_device* dev, const char* path, uint32_t row_bytes,
                                   uint32_t flags, mg_matrix** out) {
  if (!dev || !path || !out) return mg::fail(MG_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (row_bytes == 0 || row_bytes % 16 != 0)
    return mg::fail(MG_ERR_ROWS, "row width %u is not a positive multiple of 16", row_bytes);
  const int fd = open(path, O_RDONLY | O_CLOEXEC);
  if (fd < 0) return mg::fail(MG_ERR_INVALID_ARG, "cannot open %s: errno %d", path, errno);
  struct stat sb;
  if (fstat(fd, &sb) != 0 || !S_ISREG(sb.st_mode)) {
    close(fd);
    return mg::fail(MG_ERR_INVALID_ARG, "%s is not a regular file", path);
  }
  const uint64_t file_bytes = (uint64_t)sb.st_size;
  if (file_bytes == 0) {
    close(fd);
    return mg::fail(MG_ERR_ROWS, "%s is empty", path);
  }
  if (!(flags & MG_LOAD_PAD_TAIL) && file_bytes % row_bytes != 0) {
    close(fd);
    return mg::fail(MG_ERR_ROWS, "%s holds %llu bytes, not a whole number of %u-byte rows", path,
                    (unsigned long long)file_bytes, row_bytes);
  }
  const uint64_t n_rows = (file_bytes + row_bytes - 1) / row_bytes;
  mg_matrix* m = nullptr;
  int rc = mg_matrix_create(dev, n_rows, row_bytes, &m);
  if (rc) {
    close(fd);
    return rc;
  }
  rc = mg::set_device(dev);
  if (!rc) rc = stream_file(m, fd, file_bytes);
  close(fd);
  if (rc) {
    const std::string msg = mg::last_error();
    mg_matrix_destroy(m);
    return mg::fail(rc, "%s", msg.c_str());
  }
  *out = m;
  return MG_OK;
}
