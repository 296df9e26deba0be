// ccfj_mwfj.hip -- host side of the device multi-window F-J.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <vector>

#include "../../include/ccfj.h"
#include "ccfj_internal.h"
#include "fj_kernels.h"
#include "fj_mwfj.h"

using ccfj_internal::fail;
using ccfj_internal::hip_fail;

namespace {

struct Bufs {
  std::vector<void*> p;
  ~Bufs() {
    // error paths may leave work in flight that reads these buffers
    (void)hipDeviceSynchronize();
    for (void* q : p) (void)hipFree(q);
  }
  int get(size_t bytes, void** out) {
    int rc = ccfj_internal::device_alloc(out, bytes);
    if (!rc) p.push_back(*out);
    return rc;
  }
};

}  // namespace

extern "C" int ccfj_mwfj(ccfj_ctx* ctx, int itype, int func, int nr, int npts, int nwin,
                         int nf, int nc, int fstride, const double* u, const float* r,
                         const float* f, const float* c, const int* n1, const int* n2,
                         const int* delta, const int* xidx, const double* xrows, int nx,
                         double* out) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  if (itype != 0 && itype != 1) return fail(CCFJ_EINVAL, "itype must be 0 or 1, got %d", itype);
  if (func != 0 && func != 1) return fail(CCFJ_EINVAL, "func must be 0 or 1, got %d", func);
  if (nr < 1 || npts < 1 || nwin < 1 || nf < 1 || nc < 1 || fstride < 1)
    return fail(CCFJ_EINVAL, "bad sizes (nr=%d npts=%d nwin=%d nf=%d nc=%d fstride=%d)", nr, npts,
                nwin, nf, nc, fstride);
  if (!u || !r || !f || !c || !n1 || !n2 || !delta || !xidx || !out || (nx > 0 && !xrows))
    return fail(CCFJ_EINVAL, "null pointer");
  const int nbins = npts / 2 + 1;
  if ((int64_t)(nf - 1) * fstride >= nbins)
    return fail(CCFJ_EINVAL, "frequency %d x stride %d beyond the %d rFFT bins", nf, fstride, nbins);
  hipError_t e = hipSetDevice(ccfj_internal::ctx_device(ctx));
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = ccfj_internal::ctx_stream(ctx);

  fjk::MwfjDev d{};
  d.nr = nr;
  d.npts = npts;
  d.nf = nf;
  d.nc = nc;
  d.nwin = nwin;
  d.fstride = fstride;
  d.nbins = nbins;
  d.func = func;
  d.half = (nr + 31) / 32 * 32;
  d.nrhs = (2 * nwin + 31) / 32 * 32;
  const int ndesc = nwin * nr;
  // windows per rFFT batch: all of them when the tapered traces and their
  // spectra fit in 8 GiB (64 windows of 480 x 16384 take 8 GB)
  const size_t per_win = (size_t)nr * npts * 8 + (size_t)nr * nbins * 16;
  const int cw = (int)std::max<size_t>(1, std::min<size_t>(nwin, ((size_t)8 << 30) / per_win));

  Bufs bufs;
  void *du, *dr, *drr, *df, *dc, *dn1, *dn2, *ddl, *dxi, *dxr = nullptr, *dxw, *dspec, *dout;
  int rc = CCFJ_OK;
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 7, sizeof(double) * (size_t)nr * npts, &du);
  if (!rc) rc = bufs.get(sizeof(float) * nr, &dr);
  if (!rc) rc = bufs.get(sizeof(double) * nr, &drr);
  if (!rc) rc = bufs.get(sizeof(float) * nf, &df);
  if (!rc) rc = bufs.get(sizeof(float) * nc, &dc);
  if (!rc) rc = bufs.get(sizeof(int) * ndesc, &dn1);
  if (!rc) rc = bufs.get(sizeof(int) * ndesc, &dn2);
  if (!rc) rc = bufs.get(sizeof(int) * ndesc, &ddl);
  if (!rc) rc = bufs.get(sizeof(int) * ndesc, &dxi);
  if (!rc && nx > 0) rc = bufs.get(sizeof(double) * (size_t)nx * npts, &dxr);
  // the large work arrays live in the context's grow-only scratch slots
  // 2-7 (kept between calls; slots 0-1 belong to the Hilbert transform)
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 2, sizeof(double) * (size_t)cw * nr * npts, &dxw);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 3, 16 * (size_t)cw * nr * nbins, &dspec);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 4, sizeof(double) * (size_t)nwin * nc * nf, &dout);
  if (rc) return rc;
  std::vector<double> hrr(r, r + nr);
  e = hipMemcpyAsync(du, u, sizeof(double) * (size_t)nr * npts, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dr, r, sizeof(float) * nr, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(drr, hrr.data(), sizeof(double) * nr, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(df, f, sizeof(float) * nf, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dc, c, sizeof(float) * nc, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dn1, n1, sizeof(int) * ndesc, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dn2, n2, sizeof(int) * ndesc, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(ddl, delta, sizeof(int) * ndesc, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dxi, xidx, sizeof(int) * ndesc, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && nx > 0)
    e = hipMemcpyAsync(dxr, xrows, sizeof(double) * (size_t)nx * npts, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync H2D");
  d.u = (const double*)du;
  d.r = (const float*)dr;
  d.rr = (const double*)drr;
  d.f = (const float*)df;
  d.c = (const float*)dc;
  d.n1 = (const int*)dn1;
  d.n2 = (const int*)dn2;
  d.dl = (const int*)ddl;
  d.xidx = (const int*)dxi;
  d.xrows = (const double*)dxr;
  d.xw = (double*)dxw;
  d.spec = dspec;
  d.out = (double*)dout;

  void *dB = nullptr, *dT = nullptr, *dre = nullptr, *dim = nullptr, *dtmp = nullptr;
  if (itype == 0) {
    const size_t Kp = (size_t)(func ? 2 : 1) * d.half;
    if ((rc = ccfj_internal::ctx_scratch(ctx, 5, sizeof(float) * nf * Kp * d.nrhs, &dB))) return rc;
    if ((rc = ccfj_internal::ctx_scratch(ctx, 6, sizeof(float) * (size_t)nwin * nf * nc, &dT)))
      return rc;
    e = fjk::launch_fill((float*)dB, nf * Kp * d.nrhs, 0.0f, st);
    if (e != hipSuccess) return hip_fail(e, "fill kernel launch");
    d.B = (float*)dB;
    d.T = (float*)dT;
  } else {
    if ((rc = bufs.get(sizeof(float) * (size_t)nr * nf, &dre))) return rc;
    if ((rc = bufs.get(sizeof(float) * (size_t)nr * nf, &dim))) return rc;
    if ((rc = bufs.get(sizeof(float) * (size_t)nc * nf, &dtmp))) return rc;
  }

  for (int w0 = 0; w0 < nwin; w0 += cw) {
    const int cwn = std::min(cw, nwin - w0);
    if ((e = fjk::launch_window(d, w0, cwn, st)) != hipSuccess) return hip_fail(e, "window kernel");
    hipfftHandle plan = 0;
    if ((rc = ccfj_internal::cached_fft_plan(npts, cwn * nr, HIPFFT_D2Z, st, &plan))) return rc;
    if (hipfftExecD2Z(plan, (hipfftDoubleReal*)dxw, (hipfftDoubleComplex*)dspec) != HIPFFT_SUCCESS)
      return fail(CCFJ_EDEVICE, "hipfftExecD2Z failed");
    if (itype == 0) {
      if ((e = fjk::launch_mwfj_coef(d, w0, cwn, st)) != hipSuccess) return hip_fail(e, "coef kernel");
    } else {
      for (int wl = 0; wl < cwn; wl++) {
        if ((e = fjk::launch_mwfj_stage(d, wl, (float*)dre, (float*)dim, st)) != hipSuccess)
          return hip_fail(e, "stage kernel");
        rc = ccfj_transform_dev(ctx, func ? CCFJ_MODE_ABS_H : CCFJ_MODE_ABS_J, 1, 1,
                                (const float*)dre, (const float*)dim, d.r, d.f, d.c, nr, nf, nc,
                                (float*)dtmp, st);
        if (rc) return rc;
        e = fjk::launch_to_double((const float*)dtmp, (size_t)nc * nf,
                                  d.out + (size_t)(w0 + wl) * nc * nf, st);
        if (e != hipSuccess) return hip_fail(e, "convert kernel");
      }
    }
  }
  if (itype == 0) {
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    if ((rc = ccfj_internal::timing_begin(ctx, st, &ev))) return rc;
    if ((e = fjk::launch_mwfj_mfma(d, st)) != hipSuccess) return hip_fail(e, "MFMA kernel");
    ccfj_internal::timing_end(ctx, st, ev);
    if ((e = fjk::launch_mwfj_finish(d, st)) != hipSuccess) return hip_fail(e, "finish kernel");
  }
  e = hipMemcpyAsync(out, dout, sizeof(double) * (size_t)nwin * nc * nf, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync D2H");
  return CCFJ_OK;
}
