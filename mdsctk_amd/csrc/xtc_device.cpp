// Device XTC ingest (include/mdsctk_io.h: mdsctk_xtc_read_device): the file
// is read into pinned host chunks, each uploaded while the next is read; the
// frame headers are walked as the chunks arrive (the format's chain of frame
// lengths); then every frame is decoded on the GPU, one lane per frame
// (xtc_kernels.hip).  Replaces the host read loop of knn_rms.cpp:186-206 for
// callers that compute on the device.  Host-side helpers of the format are
// in xtc.cpp; this file holds the HIP part only.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/mdsctk_io.h"
#include "../../include/mdsctk_knn.h"

extern "C" void mdsctk_io_set_error(const char* msg);

namespace mdsctk {
// xtc_kernels.hip
hipError_t launch_xtc_decode(const unsigned char* file, const int64_t* off, int64_t n, int natoms,
                             float* xyz, unsigned long long* bad, hipStream_t st);
}  // namespace mdsctk

namespace {

constexpr int kMagic = 1995;

int io_fail(int code, const char* fmt, const char* a = "", long long b = 0) {
  char buf[512];
  snprintf(buf, sizeof buf, fmt, a, b);
  mdsctk_io_set_error(buf);
  return code;
}

inline int32_t be_i32(const unsigned char* p) {
  return (int32_t)(((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]);
}

}  // namespace

extern "C" {

int mdsctk_xtc_read_device(const char* path, int32_t natoms, int device, float** d_xyz,
                           int64_t* n_read) {
  if (!path || !d_xyz || !n_read || natoms <= 0)
    return io_fail(MDSCTK_ERR_ARG, "bad arguments to xtc_read_device%s%lld", "", 0);
  *d_xyz = nullptr;
  *n_read = 0;
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return io_fail(MDSCTK_ERR_IO, "cannot open %s%lld", path, 0);
  struct stat sb;
  if (::fstat(fd, &sb) != 0) {
    ::close(fd);
    return io_fail(MDSCTK_ERR_IO, "cannot size %s%lld", path, 0);
  }
  const size_t size = (size_t)sb.st_size;
  if (size == 0) {
    ::close(fd);
    return 0;
  }
  {
    // the first frame's atom count decides before any device work
    // (knn_rms.cpp:158-178 checks it before computing)
    unsigned char h0[56];
    if (size < sizeof h0 || ::pread(fd, h0, sizeof h0, 0) != (ssize_t)sizeof h0) {
      ::close(fd);
      return io_fail(MDSCTK_ERR_FORMAT, "truncated XTC frame header%s at frame %lld", "", 0);
    }
    if (be_i32(h0) != kMagic) {
      ::close(fd);
      return io_fail(MDSCTK_ERR_FORMAT, "bad XTC magic%s in frame %lld", "", 0);
    }
    if (be_i32(h0 + 4) != natoms) {
      ::close(fd);
      return io_fail(MDSCTK_ERR_NATOMS, "%s has a different number of atoms (%lld)", path,
                     be_i32(h0 + 4));
    }
  }
  // Everything the call holds, released on every exit path.
  struct Held {
    int fd = -1;
    unsigned char* d_file = nullptr;
    int64_t* d_off = nullptr;
    unsigned long long* d_bad = nullptr;
    unsigned char* pinned[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~Held() {
      if (st) (void)hipStreamSynchronize(st);
      if (fd >= 0) ::close(fd);
      (void)hipFree(d_file);
      (void)hipFree(d_off);
      (void)hipFree(d_bad);
      for (auto* p : pinned)
        if (p) (void)hipHostFree(p);
      for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
      if (st) (void)hipStreamDestroy(st);
    }
  } h;
  h.fd = fd;
  auto hip_err = [&](hipError_t e, const char* what) {
    return io_fail(e == hipErrorOutOfMemory ? MDSCTK_ERR_NOMEM : MDSCTK_ERR_HIP,
                   "xtc_read_device: %s failed%lld", what, (long long)e);
  };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  if ((e = hipMalloc((void**)&h.d_file, size + 64)) != hipSuccess) return hip_err(e, "hipMalloc");
  constexpr size_t kChunk = size_t(64) << 20;
  const size_t chunk = std::min(kChunk, (size + 4095) & ~size_t(4095));
  for (auto& p : h.pinned)
    if ((e = hipHostMalloc((void**)&p, chunk, hipHostMallocDefault)) != hipSuccess)
      return hip_err(e, "hipHostMalloc");
  if ((e = hipStreamCreateWithFlags(&h.st, hipStreamNonBlocking)) != hipSuccess)
    return hip_err(e, "hipStreamCreate");
  for (auto& ev : h.ev)
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
      return hip_err(e, "hipEventCreate");

  // Read chunk c into pinned buffer c % 2 while chunk c-1 uploads; walk the
  // frame headers of each chunk as it arrives (a header may straddle two
  // chunks: the last bytes of the previous chunk are kept in `tail`).
  std::vector<int64_t> off;
  size_t pos = 0;  // next frame start
  unsigned char tail[128];
  size_t tail_base = 0, tail_len = 0;  // tail holds bytes [tail_base, tail_base + tail_len)
  unsigned char hdr[92];
  for (size_t c0 = 0, ci = 0; c0 < size; c0 += chunk, ++ci) {
    const size_t len = std::min(chunk, size - c0);
    unsigned char* buf = h.pinned[ci & 1];
    if (ci >= 2 && (e = hipEventSynchronize(h.ev[ci & 1])) != hipSuccess)
      return hip_err(e, "hipEventSynchronize");
    for (size_t got = 0; got < len;) {
      const ssize_t r = ::pread(h.fd, buf + got, len - got, (off_t)(c0 + got));
      if (r < 0 && errno == EINTR) continue;
      if (r <= 0) return io_fail(MDSCTK_ERR_IO, "short read on %s%lld", path, 0);
      got += (size_t)r;
    }
    if ((e = hipMemcpyAsync(h.d_file + c0, buf, len, hipMemcpyHostToDevice, h.st)) != hipSuccess)
      return hip_err(e, "hipMemcpyAsync");
    if ((e = hipEventRecord(h.ev[ci & 1], h.st)) != hipSuccess) return hip_err(e, "hipEventRecord");
    const size_t c1 = c0 + len;
    // bytes [p, p + m) of the file, from the tail and this chunk
    auto fetch = [&](size_t p, size_t m) -> const unsigned char* {
      if (p >= c0) return buf + (p - c0);
      for (size_t j = 0; j < m; ++j) {
        const size_t a = p + j;
        hdr[j] = a >= c0 ? buf[a - c0] : tail[a - tail_base];
      }
      return hdr;
    };
    for (;;) {
      if (pos + 56 > c1) break;
      const unsigned char* hp = fetch(pos, 56);
      if (be_i32(hp) != kMagic)
        return io_fail(MDSCTK_ERR_FORMAT, "bad XTC magic%s in frame %lld", "", (long long)off.size());
      const int na = be_i32(hp + 4), na2 = be_i32(hp + 52);
      if (na <= 0 || na != na2)
        return io_fail(MDSCTK_ERR_FORMAT, "inconsistent XTC atom count%s in frame %lld", "",
                       (long long)off.size());
      if (na != natoms)
        return io_fail(MDSCTK_ERR_NATOMS, "%s has a different number of atoms (%lld)", path, na);
      size_t flen;
      if (na <= 9) {
        flen = 56 + (size_t)12 * na;
      } else {
        if (pos + 92 > c1) break;
        const int nbytes = be_i32(fetch(pos, 92) + 88);
        if (nbytes < 0)
          return io_fail(MDSCTK_ERR_FORMAT, "negative XTC block size%s in frame %lld", "",
                         (long long)off.size());
        flen = 92 + (((size_t)nbytes + 3) & ~size_t(3));
      }
      off.push_back((int64_t)pos);
      pos += flen;
    }
    // keep the chunk's last bytes for a header that straddles into the next
    tail_len = std::min<size_t>(sizeof tail, len);
    tail_base = c1 - tail_len;
    std::memcpy(tail, buf + (len - tail_len), tail_len);
  }
  if (pos != size)
    return io_fail(MDSCTK_ERR_FORMAT, "truncated XTC frame%s %lld", "", (long long)off.size());
  const int64_t n = (int64_t)off.size();
  off.push_back((int64_t)size);
  float* xyz = nullptr;
  if ((e = hipMalloc((void**)&h.d_off, off.size() * sizeof(int64_t))) != hipSuccess)
    return hip_err(e, "hipMalloc");
  if ((e = hipMalloc((void**)&h.d_bad, sizeof(unsigned long long))) != hipSuccess)
    return hip_err(e, "hipMalloc");
  if ((e = hipMalloc((void**)&xyz, std::max<size_t>(1, (size_t)n * natoms * 3 * sizeof(float)))) !=
      hipSuccess)
    return hip_err(e, "hipMalloc");
  unsigned long long bad = 0;
  e = hipMemcpyAsync(h.d_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                     h.st);
  if (e == hipSuccess)
    e = mdsctk::launch_xtc_decode(h.d_file, h.d_off, n, natoms, xyz, h.d_bad, h.st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&bad, h.d_bad, sizeof bad, hipMemcpyDeviceToHost, h.st);
  if (e == hipSuccess) e = hipStreamSynchronize(h.st);
  if (e != hipSuccess) {
    (void)hipFree(xyz);
    return hip_err(e, "XTC decode");
  }
  if (bad != ~0ull) {
    (void)hipFree(xyz);
    return io_fail(MDSCTK_ERR_FORMAT, "malformed XTC frame%s %lld", "", (long long)bad);
  }
  *d_xyz = xyz;
  *n_read = n;
  return 0;
}

int mdsctk_xtc_free_device(float* d_xyz) {
  return hipFree(d_xyz) == hipSuccess ? 0 : MDSCTK_ERR_HIP;
}

}  // extern "C"
