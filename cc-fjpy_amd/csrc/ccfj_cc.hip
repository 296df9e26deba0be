// ccfj_cc.hip -- host side of the device cross-correlation (fj_cc.hip).
//
// Reference: src/CrossCorr.cpp:33-159 (CrossCorrelation) and its Cython
// shim src/ccfj.pyx:44-69 (CC_full).  The window count is decided on the
// host as the reference decides it; the device builds the "station has
// this window" mask and the per-pair window counts by the reference's rule,
// and does the FFTs, the whitening, the pair products and the averaging.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>

#include "../../include/ccfj.h"
#include "ccfj_internal.h"
#include "fj_cc.h"
#include "fj_kernels.h"

using ccfj_internal::fail;
using ccfj_internal::hip_fail;

namespace {

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

int check_sizes(int nsta, int npts, int nf, int fstride, int fftlen, int steplen, int npairs) {
  if (nsta < 1 || npts < 1 || nf < 1 || fstride < 1 || fftlen < 1 || npairs < 0)
    return fail(CCFJ_EINVAL, "bad sizes (nsta=%d npts=%d nf=%d fstride=%d fftlen=%d npairs=%d)",
                nsta, npts, nf, fstride, fftlen, npairs);
  if (fftlen < npts && steplen < 1) return fail(CCFJ_EINVAL, "step length %d", steplen);
  const int len = fftlen >= npts ? npts : fftlen;
  const int nbins = len / 2 + 1;
  if ((long)(nf - 1) * fstride >= nbins)
    return fail(CCFJ_EINVAL, "bin %d x stride %d beyond the %d rFFT bins", nf - 1, fstride, nbins);
  return CCFJ_OK;
}

// The cross-correlation on device pointers, enqueued on st.  data is
// replaced by its signs when onebit is set (as the reference does to its
// caller's array).  re/im hold the caller's running sums when accumulate
// is set (the reference's +=), and receive the averages.  Work arrays live
// in the context's grow-only scratch slots 3-6 and 8 (2, 7 and 9 belong to
// the host-pointer entry point), so a repeated call allocates nothing.
int cc_run(ccfj_ctx* ctx, int nsta, int npts, int nf, int fstride, int fftlen, int steplen,
           int onebit, int whiten, int npairs, const int* pairs, const int* startend, float* data,
           float* re, float* im, bool accumulate, hipStream_t st) {
  // window count (src/CrossCorr.cpp:50-56): a window longer than the record
  // is cut to it and used once
  long nshifts;
  if (fftlen >= npts) {
    fftlen = npts;
    nshifts = 1;
  } else {
    nshifts = 1 + (long)(npts - fftlen) / steplen;
  }
  // with a bin stride that divides the window, each window is folded to
  // L = fftlen / fstride samples and its length-L rFFT gives exactly the
  // bins that are kept (every fstride-th of the length-fftlen one)
  const int fold = (fstride > 1 && fftlen % fstride == 0) ? fstride : 1;
  const int L = fftlen / fold;
  const int bstride = fold > 1 ? 1 : fstride;
  // even L: a length-L/2 complex FFT of the packed samples, the used bins
  // formed from it on read
  const int zhalf = L % 2 == 0 && L >= 2;
  const int nbins = zhalf ? L / 2 : L / 2 + 1;   // spec row length
  // windows per batch: gathered segments, their spectra and the extracted
  // bins within 4 GiB (and within one grid dimension)
  const size_t per_win = (size_t)nsta * ((size_t)L * 4 + (size_t)nbins * 8 + (size_t)nf * 8);
  const int cw = (int)std::max<size_t>(
      1, std::min<size_t>(std::min<long>(nshifts, 65535), ((size_t)4 << 30) / per_win));

  // Gram form for dense pair lists (at least 1/8 of all station pairs),
  // the per-pair form otherwise
  const bool gram = nsta <= 4096 && (nf + 3) / 4 <= 65535 &&
                    (long)npairs * 8 >= (long)nsta * (nsta - 1);
  const int ntile = (nsta + 31) / 32;
  void *dseg, *dspec, *dS, *dacc, *dmask;
  const size_t use_bytes = align256((size_t)nshifts * nsta);
  const size_t cnt_bytes = align256(sizeof(float) * npairs);
  const size_t dup_bytes = gram ? align256((size_t)npairs) : 0;
  const size_t tile_bytes = gram ? align256((size_t)ntile * ntile) : 0;
  const size_t pidx_bytes = gram ? align256(sizeof(int) * (size_t)nsta * nsta) : 0;
  const size_t flag_bytes = gram ? sizeof(int) : 0;
  int rc = ccfj_internal::ctx_scratch(ctx, 3, sizeof(float) * (size_t)cw * nsta * L, &dseg);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 4, 8 * (size_t)cw * nsta * nbins, &dspec);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 5, 8 * (size_t)cw * nsta * nf, &dS);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 6, 8 * (size_t)npairs * nf, &dacc);
  if (!rc)
    rc = ccfj_internal::ctx_scratch(
        ctx, 8, use_bytes + cnt_bytes + dup_bytes + tile_bytes + pidx_bytes + flag_bytes, &dmask);
  if (rc) return rc;
  char* m = (char*)dmask;
  unsigned char* duse = (unsigned char*)m;
  float* dcount = (float*)(m + use_bytes);
  unsigned char* ddup = (unsigned char*)(m + use_bytes + cnt_bytes);
  unsigned char* dtiles = (unsigned char*)(m + use_bytes + cnt_bytes + dup_bytes);
  int* dpidx = (int*)(m + use_bytes + cnt_bytes + dup_bytes + tile_bytes);
  int* dflag = (int*)(m + use_bytes + cnt_bytes + dup_bytes + tile_bytes + pidx_bytes);
  hipError_t e;
  if ((e = fjk::launch_cc_use(startend, nsta, (int)nshifts, steplen, fftlen, duse, st)) != hipSuccess)
    return hip_fail(e, "use kernel");
  if ((e = fjk::launch_cc_count(duse, pairs, npairs, nsta, (int)nshifts, dcount, st)) != hipSuccess)
    return hip_fail(e, "count kernel");
  if (gram) {
    e = fjk::launch_cc_fill_i32(dpidx, (size_t)nsta * nsta, -1, st);
    if (e == hipSuccess) e = fjk::launch_fill((float*)dtiles, tile_bytes / 4, 0.0f, st);
    if (e == hipSuccess)
      e = fjk::launch_cc_pair_map(pairs, npairs, nsta, ntile, dpidx, ddup, dtiles, st);
    if (e != hipSuccess) return hip_fail(e, "pair map");
  }
  if (onebit && (e = fjk::launch_cc_onebit(data, (size_t)nsta * npts, st)) != hipSuccess)
    return hip_fail(e, "one-bit kernel");
  if (accumulate) {
    if ((e = fjk::launch_cc_interleave(re, im, (size_t)npairs * nf, dacc, st)) != hipSuccess)
      return hip_fail(e, "initial sums");
  } else if (!(gram && nshifts > 0) &&   // the Gram form stores every pair on the first chunk
             (e = fjk::launch_fill((float*)dacc, 2 * (size_t)npairs * nf, 0.0f, st)) != hipSuccess) {
    return hip_fail(e, "fill kernel");
  }
  for (long w0 = 0; w0 < nshifts; w0 += cw) {
    const int n = (int)std::min<long>(cw, nshifts - w0);
    if ((e = fjk::launch_cc_gather(data, nsta, npts, L, fold, steplen, (int)w0, n, (float*)dseg,
                                   st)) != hipSuccess)
      return hip_fail(e, "gather kernel");
    hipfftHandle plan = 0;
    if (zhalf) {
      if ((rc = ccfj_internal::cached_fft_plan(ctx, L / 2, n * nsta, HIPFFT_C2C, st, &plan))) return rc;
      if (hipfftExecC2C(plan, (hipfftComplex*)dseg, (hipfftComplex*)dspec, HIPFFT_FORWARD) !=
          HIPFFT_SUCCESS)
        return fail(CCFJ_EDEVICE, "hipfftExecC2C failed");
    } else {
      if ((rc = ccfj_internal::cached_fft_plan(ctx, L, n * nsta, HIPFFT_R2C, st, &plan))) return rc;
      if (hipfftExecR2C(plan, (hipfftReal*)dseg, (hipfftComplex*)dspec) != HIPFFT_SUCCESS)
        return fail(CCFJ_EDEVICE, "hipfftExecR2C failed");
    }
    if (gram) {
      e = fjk::launch_cc_fill_i32(dflag, 1, 0, st);
      if (e == hipSuccess)
        e = fjk::launch_cc_spec_t(dspec, n * nsta, nsta, nbins, zhalf, nf, bstride, whiten, duse,
                                  (int)w0, dS, dflag, st);
      if (e == hipSuccess) {
        // the pair stage's main kernel, timed when the context's timing is on
        std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
        if ((rc = ccfj_internal::timing_begin(ctx, st, &ev))) return rc;
        e = fjk::launch_cc_gram(dS, nsta, nf, n, ntile, whiten, dtiles, dpidx, dflag,
                                w0 == 0 && !accumulate, dacc, st);
        ccfj_internal::timing_end(ctx, st, ev);
      }
      if (e == hipSuccess)
        e = fjk::launch_cc_pairs_t(dS, whiten, duse, pairs, npairs, nsta, nf, (int)w0, n, ddup,
                                   dflag, w0 == 0 && !accumulate, dacc, st);
      if (e != hipSuccess) return hip_fail(e, "Gram pair stage");
    } else {
      if ((e = fjk::launch_cc_spec(dspec, n * nsta, nbins, zhalf, nf, bstride, whiten, dS, st)) !=
          hipSuccess)
        return hip_fail(e, "spectrum kernel");
      if ((e = fjk::launch_cc_pairs(dS, duse, pairs, npairs, nsta, nf, (int)w0, n, dacc, st)) !=
          hipSuccess)
        return hip_fail(e, "pair kernel");
    }
  }
  if ((e = fjk::launch_cc_finish(dacc, dcount, npairs, nf, re, im, st)) != hipSuccess)
    return hip_fail(e, "finish kernel");
  return CCFJ_OK;
}

// ncfsi == nullptr: ncfsr receives interleaved complex [npairs][nf][2]
// (accumulate must be false then)
int cross_correlate(ccfj_ctx* ctx, int nsta, int npts, int nf, int fstride, int fftlen,
                    int steplen, int onebit, int whiten, int npairs, const int* pairs,
                    const int* startend, const float* data, float* ncfsr, float* ncfsi,
                    bool accumulate) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  auto lk = ccfj_internal::ctx_lock(ctx);
  int rc = check_sizes(nsta, npts, nf, fstride, fftlen, steplen, npairs);
  if (rc) return rc;
  const bool interleaved = ncfsi == nullptr;
  if (!startend || !data || (npairs > 0 && !pairs) || (npairs > 0 && !ncfsr) ||
      (interleaved && accumulate))
    return fail(CCFJ_EINVAL, "null pointer");
  for (int p = 0; p < 2 * npairs; p++)
    if (pairs[p] < 0 || pairs[p] >= nsta) return fail(CCFJ_EINVAL, "pair station %d of %d", pairs[p], nsta);
  if (npairs == 0) return CCFJ_OK;

  hipError_t e = hipSetDevice(ccfj_internal::ctx_device(ctx));
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = ccfj_internal::ctx_stream(ctx);
  // records in slot 2, running sums / results in slot 7, pairs and
  // start/end samples in slot 9
  void *ddata, *dres, *dlist;
  const size_t nout = (size_t)npairs * nf;
  const size_t pbytes = align256(sizeof(int) * 2 * (size_t)npairs);
  rc = ccfj_internal::ctx_scratch(ctx, 2, sizeof(float) * (size_t)nsta * npts, &ddata);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 7, 2 * align256(sizeof(float) * nout), &dres);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 9, pbytes + sizeof(int) * 2 * (size_t)nsta, &dlist);
  if (rc) return rc;
  float* dre = (float*)dres;   // interleaved: the whole slot, 2 nout floats
  float* dim = (float*)((char*)dres + align256(sizeof(float) * nout));
  int* dpairs = (int*)dlist;
  int* dse = (int*)((char*)dlist + pbytes);
  if ((rc = ccfj_internal::copy_h2d(ctx, ddata, data, sizeof(float) * (size_t)nsta * npts, st)))
    return rc;
  e = hipMemcpyAsync(dpairs, pairs, sizeof(int) * 2 * (size_t)npairs, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dse, startend, sizeof(int) * 2 * (size_t)nsta, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync H2D");
  if (accumulate) {
    if ((rc = ccfj_internal::copy_h2d(ctx, dre, ncfsr, sizeof(float) * nout, st))) return rc;
    if ((rc = ccfj_internal::copy_h2d(ctx, dim, ncfsi, sizeof(float) * nout, st))) return rc;
  }
  rc = cc_run(ctx, nsta, npts, nf, fstride, fftlen, steplen, onebit, whiten, npairs, dpairs, dse,
              (float*)ddata, dre, interleaved ? nullptr : dim, accumulate, st);
  if (rc) {
    (void)hipStreamSynchronize(st);
    return rc;
  }
  if (interleaved) return ccfj_internal::copy_d2h(ctx, ncfsr, dre, 2 * sizeof(float) * nout, st);
  if ((rc = ccfj_internal::copy_d2h(ctx, ncfsr, dre, sizeof(float) * nout, st))) return rc;
  return ccfj_internal::copy_d2h(ctx, ncfsi, dim, sizeof(float) * nout, st);
}

}  // namespace

extern "C" {

int ccfj_cross_correlate(ccfj_ctx* ctx, int nsta, int npts, int nf, int fstride, int fftlen,
                         int steplen, int onebit, int whiten, int npairs, const int* pairs,
                         const int* startend, const float* data, float* ncfsr, float* ncfsi) {
  return cross_correlate(ctx, nsta, npts, nf, fstride, fftlen, steplen, onebit, whiten, npairs,
                         pairs, startend, data, ncfsr, ncfsi, false);
}

int ccfj_cross_correlate_c64(ccfj_ctx* ctx, int nsta, int npts, int nf, int fstride, int fftlen,
                             int steplen, int onebit, int whiten, int npairs, const int* pairs,
                             const int* startend, const float* data, float* ncfs) {
  if (npairs > 0 && !ncfs) return fail(CCFJ_EINVAL, "null pointer");
  return cross_correlate(ctx, nsta, npts, nf, fstride, fftlen, steplen, onebit, whiten, npairs,
                         pairs, startend, data, ncfs, nullptr, false);
}

int ccfj_cross_correlate_dev(ccfj_ctx* ctx, int nsta, int npts, int nf, int fstride, int fftlen,
                             int steplen, int onebit, int whiten, int npairs, const int* pairs,
                             const int* startend, float* data, float* ncfsr, float* ncfsi,
                             void* stream) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  int rc = check_sizes(nsta, npts, nf, fstride, fftlen, steplen, npairs);
  if (rc) return rc;
  if (!startend || !data || (npairs > 0 && (!pairs || !ncfsr || !ncfsi)))
    return fail(CCFJ_EINVAL, "null pointer");
  if (npairs == 0) return CCFJ_OK;
  hipError_t e = hipSetDevice(ccfj_internal::ctx_device(ctx));
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = stream ? (hipStream_t)stream : ccfj_internal::ctx_stream(ctx);
  return cc_run(ctx, nsta, npts, nf, fstride, fftlen, steplen, onebit, whiten, npairs, pairs,
                startend, data, ncfsr, ncfsi, false, st);
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
