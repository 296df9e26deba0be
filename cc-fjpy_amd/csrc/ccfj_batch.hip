// ccfj_batch.hip -- host side of the batched transform: plan, autotuner,
// HIP-graph replay.
//
// A plan fixes the sub-array sizes, the grids, the mode and itype.  Its first
// run prepares the coefficients, then times the persistent kernel with each
// candidate velocity tile (64/128/256 lanes per workgroup) on the middle
// eighth of the work items and keeps the fastest.  With use_graph != 0 the
// whole launch sequence (station arrays, coefficients, counter reset,
// persistent kernel, transpose with optional normalisation) is captured
// once into a HIP graph and replayed while the device pointers stay the
// same.  The work-queue counter is reset by a kernel, not a memset, so the
// captured graph holds no memset node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "../../include/ccfj.h"
#include "ccfj_internal.h"
#include "fj_batched.h"
#include "fj_kernels.h"
#include "fj_lane.h"

using ccfj_internal::fail;
using ccfj_internal::hip_fail;

struct ccfj_batch {
  ccfj_ctx* ctx = nullptr;
  int device = 0;
  int mode = 0, itype = 0, normalize = 0;
  int nb = 0, nf = 0, nc = 0, max_nrp = 0;
  int64_t rows = 0;
  fjk::BatchDev d{};
  std::vector<void*> owned;
  int tile = 0;             // 0 until tuned
  int grid[3] = {0, 0, 0};  // persistent grid per candidate tile
  unsigned int total = 0;
  // graph cache
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* key[6] = {};
};

namespace {

constexpr int kTiles[3] = {64, 128, 256};

int tile_index(int t) { return t == 64 ? 0 : (t == 128 ? 1 : 2); }

template <typename T>
int upload(ccfj_batch* p, const std::vector<T>& h, const T** dst) {
  void* q = nullptr;
  int rc = ccfj_internal::device_alloc(&q, sizeof(T) * h.size());
  if (rc) return rc;
  p->owned.push_back(q);
  hipError_t e = hipMemcpy(q, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy H2D");
  *dst = (const T*)q;
  return CCFJ_OK;
}

int alloc(ccfj_batch* p, size_t bytes, void** dst) {
  int rc = ccfj_internal::device_alloc(dst, bytes);
  if (rc) return rc;
  p->owned.push_back(*dst);
  return CCFJ_OK;
}

void drop_graph(ccfj_batch* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

int enqueue(ccfj_batch* p, hipStream_t st, unsigned int limit, int tile) {
  hipError_t e = fjk::launch_prep_batched(p->mode, p->itype, p->d, p->max_nrp, st);
  if (e != hipSuccess) return hip_fail(e, "batched prep launch");
  const void* k = fjk::batched_kernel(p->mode, p->itype, tile);
  e = fjk::launch_batched(k, tile, p->grid[tile_index(tile)], p->d, 0, limit, st);
  if (e != hipSuccess) return hip_fail(e, "batched kernel launch");
  e = fjk::launch_finish(p->d.T, p->normalize ? p->d.m : nullptr, p->nb, p->nf, p->nc, p->d.out,
                         st);
  if (e != hipSuccess) return hip_fail(e, "batched finish launch");
  return CCFJ_OK;
}

// Time each candidate tile on the middle eighth of the work queue (the
// median-sized sub-arrays: the queue runs from the largest down).
int autotune(ccfj_batch* p, hipStream_t st) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(CCFJ_EDEVICE, "hipEventCreate failed");
  hipError_t e = fjk::launch_prep_batched(p->mode, p->itype, p->d, p->max_nrp, st);
  if (e != hipSuccess) return hip_fail(e, "batched prep launch");
  float best = 1e30f;
  int best_tile = 256;
  for (int t : kTiles) {
    const int ncb = (p->nc + t - 1) / t;
    const unsigned int total = (unsigned int)((int64_t)p->nb * p->nf * ncb);
    const unsigned int first = (unsigned int)((uint64_t)total * 7 / 16);
    const unsigned int limit = std::max(first + 1, (unsigned int)((uint64_t)total * 9 / 16));
    const void* k = fjk::batched_kernel(p->mode, p->itype, t);
    float ms = 0.f;
    for (int rep = 0; rep < 2; rep++) {  // first pass warms caches and code
      (void)hipEventRecord(e0, st);
      e = fjk::launch_batched(k, t, p->grid[tile_index(t)], p->d, first, limit, st);
      if (e != hipSuccess) return hip_fail(e, "batched kernel launch (tuning)");
      (void)hipEventRecord(e1, st);
      e = hipEventSynchronize(e1);
      if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
      (void)hipEventElapsedTime(&ms, e0, e1);
    }
    // normalise by the fraction of the work the sample covered
    const float per_total = ms * (float)total / (float)(limit - first);
    if (per_total < best) {
      best = per_total;
      best_tile = t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p->tile = best_tile;
  return CCFJ_OK;
}

}  // namespace

extern "C" {

int ccfj_batch_create(ccfj_ctx* ctx, int mode, int itype, int normalize, int nb,
                      const long long* soff, int nf, int nc, ccfj_batch** out) {
  if (!ctx || !out || !soff) return fail(CCFJ_EINVAL, "null argument");
  *out = nullptr;
  if (mode < CCFJ_MODE_J || mode > CCFJ_MODE_ABS_H) return fail(CCFJ_EINVAL, "unknown mode %d", mode);
  if (itype != 0 && itype != 1) return fail(CCFJ_EINVAL, "itype must be 0 or 1, got %d", itype);
  if (nb <= 0 || nf <= 0 || nc <= 0) return fail(CCFJ_EINVAL, "empty batch (nb=%d nf=%d nc=%d)", nb, nf, nc);
  for (int b = 0; b < nb; b++)
    if (soff[b + 1] < soff[b]) return fail(CCFJ_EINVAL, "station offsets must not decrease");
  if ((int64_t)nb * nf * ((nc + 63) / 64) >= (1LL << 32) - (1LL << 24))
    return fail(CCFJ_EINVAL, "batch too large for the 32-bit work counter");
  const int dev = ccfj_internal::ctx_device(ctx);
  hipError_t e = hipSetDevice(dev);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");

  ccfj_batch* p = new ccfj_batch();
  p->ctx = ctx;
  p->device = dev;
  p->mode = mode;
  p->itype = itype;
  p->normalize = normalize;
  p->nb = nb;
  p->nf = nf;
  p->nc = nc;
  p->rows = soff[nb];
  std::vector<int64_t> hsoff(soff, soff + nb + 1), hpoff(nb), hcoff(nb);
  std::vector<int> hnrp(nb), horder(nb);
  int64_t pacc = 0, cacc = 0;
  for (int b = 0; b < nb; b++) {
    const int nr = (int)(soff[b + 1] - soff[b]);
    hnrp[b] = nr > 0 ? fjl::padded_stations(nr) : 0;
    p->max_nrp = std::max(p->max_nrp, hnrp[b]);
    hpoff[b] = pacc;
    pacc += hnrp[b] + fjl::kUnroll;
    hcoff[b] = cacc;
    cacc += (int64_t)nf * (int64_t)fjl::coef_block(hnrp[b], mode, itype);
  }
  std::iota(horder.begin(), horder.end(), 0);
  std::stable_sort(horder.begin(), horder.end(), [&](int a, int b) { return hnrp[a] > hnrp[b]; });

  int rc = CCFJ_OK;
  void *rr, *hh, *ir, *isr, *bw, *dr, *coef, *counter, *T, *m;
  if (!rc) rc = upload(p, hsoff, &p->d.soff);
  if (!rc) rc = upload(p, hpoff, &p->d.poff);
  if (!rc) rc = upload(p, hcoff, &p->d.coff);
  if (!rc) rc = upload(p, hnrp, &p->d.nrp);
  if (!rc) rc = upload(p, horder, &p->d.order);
  if (!rc) rc = alloc(p, sizeof(double) * (pacc + 1), &rr);
  if (!rc) rc = alloc(p, sizeof(float) * (pacc + 1), &hh);
  if (!rc) rc = alloc(p, sizeof(float) * (pacc + 1), &ir);
  if (!rc) rc = alloc(p, sizeof(float) * (pacc + 1), &isr);
  if (!rc) rc = alloc(p, sizeof(uint32_t) * (pacc + 1), &bw);
  if (!rc) rc = alloc(p, sizeof(float) * (pacc + 1), &dr);
  if (!rc) rc = alloc(p, sizeof(float) * (cacc + 1), &coef);
  if (!rc) rc = alloc(p, 64, &counter);
  if (!rc) rc = alloc(p, sizeof(float) * (size_t)nb * nf * nc, &T);
  if (!rc) rc = alloc(p, sizeof(float) * (size_t)nb * nf, &m);
  if (rc) {
    ccfj_batch_destroy(p);
    return rc;
  }
  p->d.nb = nb;
  p->d.nf = nf;
  p->d.nc = nc;
  p->d.rr = (double*)rr;
  p->d.hh = (float*)hh;
  p->d.ir = (float*)ir;
  p->d.isr = (float*)isr;
  p->d.bw = (uint32_t*)bw;
  p->d.dr = (float*)dr;
  p->d.coef = (float*)coef;
  p->d.counter = (unsigned int*)counter;
  p->d.T = (float*)T;
  p->d.m = (float*)m;

  int cus = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) {
    ccfj_batch_destroy(p);
    return hip_fail(e, "hipDeviceGetAttribute");
  }
  for (int t : kTiles) {
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fjk::batched_kernel(mode, itype, t), t, 0);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    p->grid[tile_index(t)] = per_cu * cus;
  }
  *out = p;
  return CCFJ_OK;
}

int ccfj_batch_run(ccfj_batch* p, const float* U0, const float* U1, const float* r,
                   const float* f, const float* c, float* out, void* stream, int use_graph) {
  if (!p) return fail(CCFJ_EINVAL, "null plan");
  const bool two = !(p->mode == CCFJ_MODE_J || p->mode == CCFJ_MODE_Y);
  if (!U0 || !r || !f || !c || !out || (two && !U1)) return fail(CCFJ_EINVAL, "null device pointer");
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = stream ? (hipStream_t)stream : ccfj_internal::ctx_stream(p->ctx);
  const void* key[6] = {U0, two ? U1 : U0, r, f, c, out};
  const bool same = std::equal(key, key + 6, p->key);
  if (!same) {
    drop_graph(p);
    std::copy(key, key + 6, p->key);
  }
  p->d.U0 = U0;
  p->d.U1 = two ? U1 : U0;
  p->d.r = r;
  p->d.f = f;
  p->d.c = c;
  p->d.out = out;
  int rc;
  if (!p->tile && (rc = autotune(p, st))) return rc;
  const int ncb = (p->nc + p->tile - 1) / p->tile;
  const unsigned int total = (unsigned int)((int64_t)p->nb * p->nf * ncb);
  if (!use_graph) return enqueue(p, st, total, p->tile);
  if (!p->exec) {
    if (st == nullptr) return fail(CCFJ_EINVAL, "graph capture needs a non-null stream");
    e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    rc = enqueue(p, st, total, p->tile);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    p->graph = g;
    e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(p);
      return hip_fail(e, "hipGraphInstantiate");
    }
  }
  e = hipGraphLaunch(p->exec, st);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "hipGraphLaunch");
}

int ccfj_batch_tile(ccfj_batch* p) { return p ? p->tile : 0; }

int ccfj_batch_destroy(ccfj_batch* p) {
  if (!p) return CCFJ_OK;
  (void)hipSetDevice(p->device);
  (void)hipDeviceSynchronize();
  drop_graph(p);
  for (void* q : p->owned) (void)hipFree(q);
  delete p;
  return CCFJ_OK;
}

int ccfj_transform_batched(ccfj_ctx* ctx, int mode, int itype, int normalize, int nb,
                           const long long* soff, const float* U0, const float* U1,
                           const float* r, const float* f, const float* c, int nf, int nc,
                           float* out) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  auto lk = ccfj_internal::ctx_lock(ctx);
  ccfj_batch* p = nullptr;
  int rc = ccfj_batch_create(ctx, mode, itype, normalize, nb, soff, nf, nc, &p);
  if (rc) return rc;
  const bool two = !(mode == CCFJ_MODE_J || mode == CCFJ_MODE_Y);
  const size_t rows = (size_t)soff[nb];
  const size_t nu = rows * nf;
  void *dU0 = nullptr, *dU1 = nullptr, *dr = nullptr, *df = nullptr, *dc = nullptr, *dout = nullptr;
  hipStream_t st = ccfj_internal::ctx_stream(ctx);
  hipError_t e = hipSuccess;
  if (!rc) rc = alloc(p, sizeof(float) * nu + 4, &dU0);
  if (!rc && two) rc = alloc(p, sizeof(float) * nu + 4, &dU1);
  if (!rc) rc = alloc(p, sizeof(float) * rows + 4, &dr);
  if (!rc) rc = alloc(p, sizeof(float) * nf, &df);
  if (!rc) rc = alloc(p, sizeof(float) * nc, &dc);
  if (!rc) rc = alloc(p, sizeof(float) * (size_t)nb * nc * nf, &dout);
  if (!rc) {
    e = hipMemcpyAsync(dU0, U0, sizeof(float) * nu, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && two) e = hipMemcpyAsync(dU1, U1, sizeof(float) * nu, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, r, sizeof(float) * rows, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(df, f, sizeof(float) * nf, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dc, c, sizeof(float) * nc, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) rc = hip_fail(e, "hipMemcpyAsync H2D");
  }
  if (!rc) rc = ccfj_batch_run(p, (const float*)dU0, (const float*)dU1, (const float*)dr,
                               (const float*)df, (const float*)dc, (float*)dout, st, 0);
  if (!rc) {
    e = hipMemcpyAsync(out, dout, sizeof(float) * (size_t)nb * nc * nf, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = hip_fail(e, "hipMemcpyAsync D2H");
  }
  ccfj_batch_destroy(p);
  return rc;
}

}  // extern "C"
