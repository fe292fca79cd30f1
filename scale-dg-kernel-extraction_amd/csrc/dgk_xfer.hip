// Host <-> device copies of whole fields through pinned staging buffers.
//
// A copy from pageable host memory runs at about 10 GB/s on the MI355X
// boxes (the runtime stages it through its own small pinned buffers, one
// chunk at a time).  Here a field is cut into chunks of kChunk bytes: the
// host copy of chunk i into one of two pinned buffers is split between
// several threads, while the DMA of chunk i-1 out of the other buffer runs on
// the caller's stream.  Copies below kDirect bytes go straight through
// hipMemcpyAsync.  The two buffers (2 x 64 MiB of pinned host memory per
// device) are allocated on the first large copy and held until the process
// ends; calls are serialised by a mutex.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "dgk_internal.hpp"

namespace dgk {
namespace {

constexpr size_t kChunk = size_t(64) << 20;
constexpr size_t kDirect = size_t(16) << 20;
constexpr int kMaxDevices = 64;
constexpr int kMaxThreads = 8;

struct Staging {
  void *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ready = false;
};

std::mutex g_mu;
Staging g_staging[kMaxDevices];

// The threads of one host copy: half the CPUs this process may use, at most
// kMaxThreads (the other half keeps the DMA submission and the caller going).
int copy_threads() {
  static const int n = [] {
    const unsigned hw = std::thread::hardware_concurrency();
    return std::max(1, std::min<int>(kMaxThreads, (int)hw / 2));
  }();
  return n;
}

// Runs f(begin, end) over [0, n) split into copy_threads() contiguous parts
// (one part, on the calling thread, below `serial` items).  A part whose
// thread cannot be started runs on the calling thread.
template <class F>
void par_for(size_t n, size_t serial, F f) {
  const int T = copy_threads();
  if (T <= 1 || n < serial) {
    f(size_t(0), n);
    return;
  }
  const size_t part = (n + T - 1) / T;
  std::vector<std::thread> th;
  th.reserve(T);
  for (int t = 1; t < T; ++t) {
    const size_t b = part * t;
    if (b >= n) break;
    try {
      th.emplace_back([=] { f(b, std::min(n, b + part)); });
    } catch (...) {
      f(b, std::min(n, b + part));
    }
  }
  f(size_t(0), std::min(n, part));
  for (auto &x : th) x.join();
}

// The staging buffers of the current device, allocated on first use.
int staging(Staging **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(DGK_ERR_HIP, "staged copy: no current device");
  Staging &S = g_staging[dev];
  if (!S.ready) {
    for (int b = 0; b < 2; ++b) {
      if (!S.buf[b] && hipHostMalloc(&S.buf[b], kChunk, hipHostMallocDefault) != hipSuccess) {
        S.buf[b] = nullptr;
        return fail(DGK_ERR_NOMEM, "staged copy: hipHostMalloc of %zu bytes failed", kChunk);
      }
      if (!S.ev[b] && hipEventCreateWithFlags(&S.ev[b], hipEventDisableTiming) != hipSuccess) {
        S.ev[b] = nullptr;
        return fail(DGK_ERR_HIP, "staged copy: hipEventCreate failed");
      }
    }
    S.ready = true;
  }
  *out = &S;
  return DGK_OK;
}

#define XOK(expr, what)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return fail(DGK_ERR_HIP, "%s: %s", what, hipGetErrorString(e_)); \
  } while (0)

// memcpy of n bytes split between the copy threads (parts of whole pages).
void par_memcpy(char *dst, const char *src, size_t n) {
  const size_t pages = (n + 4095) / 4096;
  par_for(pages, 1024, [=](size_t b, size_t e) {
    const size_t lo = b * 4096, hi = std::min(n, e * 4096);
    if (lo < hi) std::memcpy(dst + lo, src + lo, hi - lo);
  });
}

// The chunk loop of a staged host-to-device copy.
int h2d_chunks(Staging *S, char *d, const char *s, size_t bytes, hipStream_t st) {
  for (size_t off = 0, i = 0; off < bytes; off += kChunk, ++i) {
    const size_t n = std::min(kChunk, bytes - off);
    const int b = (int)(i & 1);
    // the DMA out of this buffer two chunks ago has finished
    if (i >= 2) XOK(hipEventSynchronize(S->ev[b]), "hipEventSynchronize");
    par_memcpy(static_cast<char *>(S->buf[b]), s + off, n);
    XOK(hipMemcpyAsync(d + off, S->buf[b], n, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    XOK(hipEventRecord(S->ev[b], st), "hipEventRecord");
  }
  return DGK_OK;
}

// The chunk loop of a staged device-to-host copy: chunk i's DMA into buffer
// i % 2 runs while chunk i-1 is copied out of the other buffer; buffer i % 2
// was emptied at step i-1.
int d2h_chunks(Staging *S, char *d, const char *s, size_t bytes, hipStream_t st) {
  const size_t nchunk = (bytes + kChunk - 1) / kChunk;
  for (size_t i = 0; i <= nchunk; ++i) {
    if (i < nchunk) {
      const size_t off = i * kChunk, n = std::min(kChunk, bytes - off);
      XOK(hipMemcpyAsync(S->buf[i & 1], s + off, n, hipMemcpyDeviceToHost, st),
          "hipMemcpyAsync");
      XOK(hipEventRecord(S->ev[i & 1], st), "hipEventRecord");
    }
    if (i >= 1) {
      const size_t j = i - 1, off = j * kChunk, n = std::min(kChunk, bytes - off);
      XOK(hipEventSynchronize(S->ev[j & 1]), "hipEventSynchronize");
      par_memcpy(d + off, static_cast<const char *>(S->buf[j & 1]), n);
    }
  }
  return DGK_OK;
}

}  // namespace

int find_not_equal(const double *a, size_t n, double v, size_t *where) {
  std::mutex mu;
  size_t found = n;
  par_for(n, size_t(1) << 20, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i)
      if (a[i] != v) {
        std::lock_guard<std::mutex> lock(mu);
        found = std::min(found, i);
        return;
      }
  });
  *where = found;
  return found < n;
}

int copy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (bytes < kDirect) {
    XOK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    XOK(hipStreamSynchronize(st), "hipStreamSynchronize");
    return DGK_OK;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  Staging *S = nullptr;
  int rc = staging(&S);
  if (rc) return rc;
  rc = h2d_chunks(S, static_cast<char *>(dst), static_cast<const char *>(src), bytes, st);
  // every DMA out of the staging buffers has ended before they are reused,
  // also when the loop stopped at an error
  const hipError_t e = hipStreamSynchronize(st);
  if (!rc && e != hipSuccess)
    rc = fail(DGK_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
  return rc;
}

int copy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (bytes < kDirect) {
    XOK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    XOK(hipStreamSynchronize(st), "hipStreamSynchronize");
    return DGK_OK;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  Staging *S = nullptr;
  int rc = staging(&S);
  if (rc) return rc;
  rc = d2h_chunks(S, static_cast<char *>(dst), static_cast<const char *>(src), bytes, st);
  // no DMA into the staging buffers is left running, also after an error
  if (rc) (void)hipStreamSynchronize(st);
  return rc;
}

}  // namespace dgk
