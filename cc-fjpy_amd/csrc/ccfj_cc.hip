// ccfj_cc.hip -- host side of the device cross-correlation (fj_cc.hip).
//
// Reference: src/CrossCorr.cpp:33-159 (CrossCorrelation) and its Cython
// shim src/ccfj.pyx:44-69 (CC_full).  The window count, the "station has
// this window" rule and the per-pair counts are decided on the host exactly
// as the reference decides them; the device does the FFTs, the whitening,
// the pair products and the averaging.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/ccfj.h"
#include "ccfj_internal.h"
#include "fj_cc.h"
#include "fj_kernels.h"

using ccfj_internal::fail;
using ccfj_internal::hip_fail;

namespace {

struct Bufs {
  std::vector<void*> p;
  ~Bufs() {
    (void)hipDeviceSynchronize();
    for (void* q : p) (void)hipFree(q);
  }
  int get(size_t bytes, void** out) {
    int rc = ccfj_internal::device_alloc(out, bytes);
    if (!rc) p.push_back(*out);
    return rc;
  }
};

// accumulate: the sums start from the caller's ncfsr/ncfsi (as the
// reference's do; its wrapper passes zeros), else from zero
int cross_correlate(ccfj_ctx* ctx, int nsta, int npts, int nf, int fstride, int fftlen,
                    int steplen, int onebit, int whiten, int npairs, const int* pairs,
                    const int* startend, const float* data, float* ncfsr, float* ncfsi,
                    bool accumulate) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  if (nsta < 1 || npts < 1 || nf < 1 || fstride < 1 || fftlen < 1 || npairs < 0)
    return fail(CCFJ_EINVAL, "bad sizes (nsta=%d npts=%d nf=%d fstride=%d fftlen=%d npairs=%d)",
                nsta, npts, nf, fstride, fftlen, npairs);
  if (!startend || !data || (npairs > 0 && (!pairs || !ncfsr || !ncfsi)))
    return fail(CCFJ_EINVAL, "null pointer");
  // window count (src/CrossCorr.cpp:50-56): a window longer than the record
  // is cut to it and used once
  long nshifts;
  if (fftlen >= npts) {
    fftlen = npts;
    nshifts = 1;
  } else {
    if (steplen < 1) return fail(CCFJ_EINVAL, "step length %d", steplen);
    nshifts = 1 + (long)(npts - fftlen) / steplen;
  }
  const int nbins = fftlen / 2 + 1;
  if ((long)(nf - 1) * fstride >= nbins)
    return fail(CCFJ_EINVAL, "bin %d x stride %d beyond the %d rFFT bins", nf - 1, fstride, nbins);
  for (int p = 0; p < 2 * npairs; p++)
    if (pairs[p] < 0 || pairs[p] >= nsta) return fail(CCFJ_EINVAL, "pair station %d of %d", pairs[p], nsta);
  if (npairs == 0) return CCFJ_OK;

  // station i has window k when k*steplen >= start_i and
  // k*steplen + fftlen < end_i (src/CrossCorr.cpp:91-98)
  std::vector<unsigned char> use((size_t)nshifts * nsta);
  for (long k = 0; k < nshifts; k++)
    for (int i = 0; i < nsta; i++) {
      const long t0 = k * (long)steplen;
      use[(size_t)k * nsta + i] = (t0 >= startend[2 * i] && t0 + fftlen < startend[2 * i + 1]);
    }
  std::vector<float> count(npairs, 0.0f);
  for (int p = 0; p < npairs; p++) {
    long n = 0;
    for (long k = 0; k < nshifts; k++)
      n += use[(size_t)k * nsta + pairs[2 * p]] & use[(size_t)k * nsta + pairs[2 * p + 1]];
    count[p] = (float)n;   // the reference divides the float sums by it
  }

  hipError_t e = hipSetDevice(ccfj_internal::ctx_device(ctx));
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = ccfj_internal::ctx_stream(ctx);
  // windows per batch: gathered segments, their spectra and the extracted
  // bins within 4 GiB
  const size_t per_win = (size_t)nsta * ((size_t)fftlen * 4 + (size_t)nbins * 8 + (size_t)nf * 8);
  const int cw = (int)std::max<size_t>(1, std::min<size_t>(nshifts, ((size_t)4 << 30) / per_win));

  Bufs bufs;
  void *ddata, *dseg, *dspec, *dS, *dacc, *duse, *dpairs, *dcount, *dre, *dim;
  int rc = ccfj_internal::ctx_scratch(ctx, 2, sizeof(float) * (size_t)nsta * npts, &ddata);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 3, sizeof(float) * (size_t)cw * nsta * fftlen, &dseg);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 4, 8 * (size_t)cw * nsta * nbins, &dspec);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 5, 8 * (size_t)cw * nsta * nf, &dS);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 6, 8 * (size_t)npairs * nf, &dacc);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 7, 8 * (size_t)npairs * nf, &dre);
  if (!rc) rc = bufs.get(use.size(), &duse);
  if (!rc) rc = bufs.get(sizeof(int) * 2 * (size_t)npairs, &dpairs);
  if (!rc) rc = bufs.get(sizeof(float) * npairs, &dcount);
  if (rc) return rc;
  dim = (float*)dre + (size_t)npairs * nf;
  e = hipMemcpyAsync(ddata, data, sizeof(float) * (size_t)nsta * npts, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(duse, use.data(), use.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dpairs, pairs, sizeof(int) * 2 * (size_t)npairs, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dcount, count.data(), sizeof(float) * npairs, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync H2D");
  if (onebit && (e = fjk::launch_cc_onebit((float*)ddata, (size_t)nsta * npts, st)) != hipSuccess)
    return hip_fail(e, "one-bit kernel");
  if (accumulate) {
    // interleave the caller's running sums into acc (complex pairs)
    e = hipMemcpyAsync(dre, ncfsr, sizeof(float) * (size_t)npairs * nf, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dim, ncfsi, sizeof(float) * (size_t)npairs * nf, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = fjk::launch_cc_interleave((const float*)dre, (const float*)dim, (size_t)npairs * nf, dacc,
                                    st);
    if (e != hipSuccess) return hip_fail(e, "initial sums");
  } else if ((e = fjk::launch_fill((float*)dacc, 2 * (size_t)npairs * nf, 0.0f, st)) != hipSuccess) {
    return hip_fail(e, "fill kernel");
  }
  for (long w0 = 0; w0 < nshifts; w0 += cw) {
    const int n = (int)std::min<long>(cw, nshifts - w0);
    if ((e = fjk::launch_cc_gather((const float*)ddata, nsta, npts, fftlen, steplen, (int)w0, n,
                                   (float*)dseg, st)) != hipSuccess)
      return hip_fail(e, "gather kernel");
    hipfftHandle plan = 0;
    if ((rc = ccfj_internal::cached_fft_plan(fftlen, n * nsta, HIPFFT_R2C, st, &plan))) return rc;
    if (hipfftExecR2C(plan, (hipfftReal*)dseg, (hipfftComplex*)dspec) != HIPFFT_SUCCESS)
      return fail(CCFJ_EDEVICE, "hipfftExecR2C failed");
    if ((e = fjk::launch_cc_spec(dspec, n * nsta, nbins, nf, fstride, whiten, dS, st)) != hipSuccess)
      return hip_fail(e, "spectrum kernel");
    if ((e = fjk::launch_cc_pairs(dS, (const unsigned char*)duse, (const int*)dpairs, npairs, nsta,
                                  nf, (int)w0, n, dacc, st)) != hipSuccess)
      return hip_fail(e, "pair kernel");
  }
  if ((e = fjk::launch_cc_finish(dacc, (const float*)dcount, npairs, nf, (float*)dre, (float*)dim,
                                 st)) != hipSuccess)
    return hip_fail(e, "finish kernel");
  e = hipMemcpyAsync(ncfsr, dre, sizeof(float) * (size_t)npairs * nf, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ncfsi, dim, sizeof(float) * (size_t)npairs * nf, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? CCFJ_OK : hip_fail(e, "hipMemcpyAsync D2H");
}

}  // namespace

extern "C" {

int ccfj_cross_correlate(ccfj_ctx* ctx, int nsta, int npts, int nf, int fstride, int fftlen,
                         int steplen, int onebit, int whiten, int npairs, const int* pairs,
                         const int* startend, const float* data, float* ncfsr, float* ncfsi) {
  return cross_correlate(ctx, nsta, npts, nf, fstride, fftlen, steplen, onebit, whiten, npairs,
                         pairs, startend, data, ncfsr, ncfsi, false);
}

int CrossCorrelation(CC_data d) {
  ccfj_ctx* ctx = nullptr;
  int rc = ccfj_internal::default_context(&ctx);
  if (rc) return rc;
  rc = cross_correlate(ctx, d.nsta, d.npts, d.nf, d.fstride, d.fftlen, d.steplen, d.ifonebit,
                       d.ifspecwhitenning, d.npairs, d.Pairs, d.startend, d.data, d.ncfsr, d.ncfsi,
                       true);
  if (!rc && d.ifonebit) {
    // the reference replaces the caller's samples by their signs in place
    // (src/CrossCorr.cpp:76-85)
    for (size_t i = 0; i < (size_t)d.nsta * d.npts; i++) {
      const float v = d.data[i];
      d.data[i] = v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f);
    }
  }
  return rc;
}

}  // extern "C"
