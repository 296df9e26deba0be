// ccfj_stage.hip -- large host <-> device copies through a pinned ring.
//
// A pageable copy of a few hundred MB is slow twice over: the runtime stages
// it through its own small pinned buffers, and a freshly allocated
// destination (numpy's output arrays) is page-faulted by one thread.  Here
// the context keeps two pinned chunks; DMA of one chunk overlaps the host
// copy of the other, and the host copy is split over worker threads, so the
// first-touch faults of the destination are taken in parallel too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "ccfj_internal.h"

using ccfj_internal::hip_fail;

namespace {

constexpr size_t kStageMin = (size_t)16 << 20;   // smaller copies go direct

int worker_count() {
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(16u, hc ? hc : 1u));
}

struct Ring {
  std::mutex mu;
  std::condition_variable cv;
  long stage = 0;             // chunks whose device side is complete
  std::vector<int> host;      // per chunk: workers done with their slice
  bool abort = false;
};

// slice t of T of a chunk of len bytes
void slice(size_t len, int t, int T, size_t* off, size_t* n) {
  const size_t per = (len + T - 1) / T;
  *off = std::min(len, per * t);
  *n = std::min(len, *off + per) - *off;
}

}  // namespace

namespace ccfj_internal {

int copy_d2h(ccfj_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t st) {
  hipError_t e;
  if (bytes < kStageMin) {
    e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? CCFJ_OK : hip_fail(e, "hipMemcpyAsync D2H");
  }
  void* pin[2];
  hipEvent_t ev[2];
  size_t C = 0;
  int rc = ctx_pinned_ring(ctx, pin, ev, &C);
  if (rc) return rc;
  const long n = (long)((bytes + C - 1) / C);
  const int T = worker_count();
  Ring g;
  g.host.assign(n, 0);
  auto len = [&](long k) { return std::min(C, bytes - (size_t)k * C); };
  std::vector<std::thread> w;
  for (int t = 0; t < T; t++)
    w.emplace_back([&, t] {
      for (long k = 0; k < n; k++) {
        {
          std::unique_lock<std::mutex> lk(g.mu);
          g.cv.wait(lk, [&] { return g.abort || g.stage > k; });
          if (g.abort) return;
        }
        size_t off, m;
        slice(len(k), t, T, &off, &m);
        if (m) memcpy((char*)dst + (size_t)k * C + off, (const char*)pin[k & 1] + off, m);
        {
          std::lock_guard<std::mutex> lk(g.mu);
          g.host[k]++;
        }
        g.cv.notify_all();
      }
    });
  e = hipSuccess;
  for (long k = 0; k < n && e == hipSuccess; k++) {
    if (k >= 2) {   // the buffer of chunk k-2 must have been copied out
      std::unique_lock<std::mutex> lk(g.mu);
      g.cv.wait(lk, [&] { return g.host[k - 2] == T; });
    }
    e = hipMemcpyAsync(pin[k & 1], (const char*)src + (size_t)k * C, len(k),
                       hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(ev[k & 1], st);
    if (e == hipSuccess && k >= 1) e = hipEventSynchronize(ev[(k - 1) & 1]);
    if (e == hipSuccess && k >= 1) {
      std::lock_guard<std::mutex> lk(g.mu);
      g.stage = k;
    }
    g.cv.notify_all();
  }
  if (e == hipSuccess) e = hipEventSynchronize(ev[(n - 1) & 1]);
  {
    std::lock_guard<std::mutex> lk(g.mu);
    if (e == hipSuccess) g.stage = n;
    else g.abort = true;
  }
  g.cv.notify_all();
  for (auto& th : w) th.join();
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(st);   // no DMA may still touch the ring
    return hip_fail(e, "staged D2H copy");
  }
  return CCFJ_OK;
}

int copy_h2d(ccfj_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t st) {
  hipError_t e;
  if (bytes < kStageMin) {   // a pageable source is consumed before this returns
    e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    return e == hipSuccess ? CCFJ_OK : hip_fail(e, "hipMemcpyAsync H2D");
  }
  void* pin[2];
  hipEvent_t ev[2];
  size_t C = 0;
  int rc = ctx_pinned_ring(ctx, pin, ev, &C);
  if (rc) return rc;
  const long n = (long)((bytes + C - 1) / C);
  const int T = worker_count();
  Ring g;   // stage = chunks whose DMA is complete
  g.host.assign(n, 0);
  auto len = [&](long k) { return std::min(C, bytes - (size_t)k * C); };
  std::vector<std::thread> w;
  for (int t = 0; t < T; t++)
    w.emplace_back([&, t] {
      for (long k = 0; k < n; k++) {
        {   // the DMA of chunk k-2, from the same buffer, must be complete
          std::unique_lock<std::mutex> lk(g.mu);
          g.cv.wait(lk, [&] { return g.abort || g.stage >= k - 1; });
          if (g.abort) return;
        }
        size_t off, m;
        slice(len(k), t, T, &off, &m);
        if (m) memcpy((char*)pin[k & 1] + off, (const char*)src + (size_t)k * C + off, m);
        {
          std::lock_guard<std::mutex> lk(g.mu);
          g.host[k]++;
        }
        g.cv.notify_all();
      }
    });
  e = hipSuccess;
  for (long k = 0; k < n && e == hipSuccess; k++) {
    {
      std::unique_lock<std::mutex> lk(g.mu);
      g.cv.wait(lk, [&] { return g.host[k] == T; });
    }
    e = hipMemcpyAsync((char*)dst + (size_t)k * C, pin[k & 1], len(k), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(ev[k & 1], st);
    if (e == hipSuccess && k >= 1) e = hipEventSynchronize(ev[(k - 1) & 1]);
    if (e == hipSuccess && k >= 1) {
      std::lock_guard<std::mutex> lk(g.mu);
      g.stage = k;   // chunks 0 .. k-1 are on the device
    }
    g.cv.notify_all();
  }
  if (e == hipSuccess) e = hipEventSynchronize(ev[(n - 1) & 1]);
  {
    std::lock_guard<std::mutex> lk(g.mu);
    if (e != hipSuccess) g.abort = true;
  }
  g.cv.notify_all();
  for (auto& th : w) th.join();
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(st);   // no DMA may still touch the ring
    return hip_fail(e, "staged H2D copy");
  }
  return CCFJ_OK;
}

}  // namespace ccfj_internal
