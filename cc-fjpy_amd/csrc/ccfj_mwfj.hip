// ccfj_mwfj.hip -- host side of the device multi-window F-J.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <utility>

#include "../../include/ccfj.h"
#include "ccfj_internal.h"
#include "fj_kernels.h"
#include "fj_mwfj.h"

using ccfj_internal::fail;
using ccfj_internal::hip_fail;

namespace {

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

int check_sizes(int itype, int func, int nr, int npts, int nwin, int nf, int nc, int fstride,
                int nx) {
  if (itype != 0 && itype != 1) return fail(CCFJ_EINVAL, "itype must be 0 or 1, got %d", itype);
  if (func != 0 && func != 1) return fail(CCFJ_EINVAL, "func must be 0 or 1, got %d", func);
  if (nr < 1 || npts < 1 || nwin < 1 || nf < 1 || nc < 1 || fstride < 1 || nx < 0)
    return fail(CCFJ_EINVAL, "bad sizes (nr=%d npts=%d nwin=%d nf=%d nc=%d fstride=%d nx=%d)", nr,
                npts, nwin, nf, nc, fstride, nx);
  const int nbins = npts / 2 + 1;
  if ((int64_t)(nf - 1) * fstride >= nbins)
    return fail(CCFJ_EINVAL, "frequency %d x stride %d beyond the %d rFFT bins", nf, fstride, nbins);
  return CCFJ_OK;
}

// The whole multi-window transform on device pointers, enqueued on st.
// Work arrays live in the context's grow-only scratch slots 2-7 (slots 0-1
// belong to the Hilbert transform, 8-9 to the host-pointer entry point), so
// a call allocates nothing once the slots have grown.
// Sum of the windows' support lengths (k_mwfj_goertzel), from host copies
// of the descriptions.
long long support_sum(int nr, int npts, int nwin, const int* n1, const int* n2,
                      const int* xidx, int nx) {
  long long sum = 0;
  for (int d = 0; d < nwin * nr; d++) {
    if (xidx[d] >= 0 && xidx[d] < nx) {
      sum += npts;
    } else if (xidx[d] == -1) {
      const int a = std::max(n1[d], 0), b = std::min(n2[d], npts);
      if (b > a) sum += b - a;
    }
  }
  return sum;
}

// The used bins come from the Goertzel recurrence over each window's
// support when that costs less than writing and transforming the
// zero-padded rows: 2 double operations per (support sample, bin) against
// the FFT path's measured cost, 38 of them per row sample (window kernel
// and length-8192 complex FFT: 2.54 ms for 64 x 480 rows of 16,384 samples
// on the MI355X).
// CCFJ_MWFJ_SPECTRA=fft or =goertzel in the environment forces one path (a
// test and tuning switch; both give the same spectra to rounding).
bool use_goertzel(long long support, int nf, int nwin, int nr, int npts) {
  const char* force = getenv("CCFJ_MWFJ_SPECTRA");
  if (force && strcmp(force, "fft") == 0) return false;
  if (force && strcmp(force, "goertzel") == 0) return true;
  return (double)support * nf <= 38.0 * (double)nwin * nr * npts;
}

// support: the sum of the windows' support lengths, or -1 to sum them on
// the device (one wait for the stream).
int mwfj_run(ccfj_ctx* ctx, int itype, int func, int nr, int npts, int nwin, int nf, int nc,
             int fstride, const double* u, const float* r, const float* f, const float* c,
             const int* n1, const int* n2, const int* delta, const int* xidx,
             const double* xrows, int nx, long long support, double* out, hipStream_t st) {
  fjk::MwfjDev d{};
  d.nr = nr;
  d.npts = npts;
  d.nf = nf;
  d.nc = nc;
  d.nwin = nwin;
  d.fstride = fstride;
  // even trace length: a half-length complex FFT of the packed samples,
  // the used bins formed from it by the coefficient kernel (no pass over
  // the npts/2 + 1 bins of a full rFFT)
  d.zhalf = npts % 2 == 0 && npts >= 2;
  d.nbins = d.zhalf ? npts / 2 : npts / 2 + 1;
  d.func = func;
  d.half = (nr + 31) / 32 * 32;
  d.nrhs = (2 * nwin + 31) / 32 * 32;
  void *dxw = nullptr, *dspec = nullptr, *drr = nullptr;
  const size_t o_sum = align256(12 * (size_t)d.half);
  const size_t o_bad = o_sum + 256;
  const size_t o_tw = o_bad + align256(sizeof(int) * (size_t)nr);
  int rc = ccfj_internal::ctx_scratch(ctx, 7, o_tw + 16 * (size_t)npts, &drr);
  if (rc) return rc;
  d.r = r;
  hipError_t e = fjk::launch_mwfj_stations(d, (double*)drr, (float*)((double*)drr + d.half), st);
  if (e != hipSuccess) return hip_fail(e, "station kernel");
  d.u = u;
  d.n1 = n1;
  d.n2 = n2;
  d.dl = delta;
  d.xidx = xidx;
  d.xrows = xrows;
  d.nx = xrows ? nx : 0;
  if (support < 0) {
    unsigned long long* dsum = (unsigned long long*)((char*)drr + o_sum);
    unsigned long long hsum = 0;
    if ((e = fjk::launch_fill((float*)dsum, 2, 0.0f, st)) != hipSuccess)
      return hip_fail(e, "fill kernel launch");
    if ((e = fjk::launch_mwfj_support(d, dsum, st)) != hipSuccess)
      return hip_fail(e, "support kernel");
    if ((e = hipMemcpyAsync(&hsum, dsum, sizeof(hsum), hipMemcpyDeviceToHost, st)) != hipSuccess)
      return hip_fail(e, "support copy");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "support wait");
    support = (long long)hsum;
  }
  const bool direct = use_goertzel(support, nf, nwin, nr, npts);
  // windows per batch: all of them when the tapered traces and their
  // spectra fit in 8 GiB (64 windows of 480 x 16384 take 8 GB); the
  // Goertzel path keeps only the nf used bins of each row
  const size_t row_traces = direct ? 0 : (size_t)nr * npts * 8;   // tapered traces
  const size_t row_spec = (size_t)nr * (direct ? nf : d.nbins) * 16;  // spectra
  const int cw = (int)std::max<size_t>(
      1, std::min<size_t>(nwin, ((size_t)8 << 30) / (row_traces + row_spec)));
  if (!direct) rc = ccfj_internal::ctx_scratch(ctx, 2, row_traces * cw, &dxw);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 3, row_spec * cw, &dspec);
  if (rc) return rc;
  d.rr = (const double*)drr;
  d.ir = (const float*)((const double*)drr + d.half);
  d.f = f;
  d.c = c;
  d.xw = (double*)dxw;
  d.spec = dspec;
  d.out = out;

  float *dre = nullptr, *dim = nullptr, *dtmp = nullptr;
  if (itype == 0) {
    const size_t Kp = (size_t)(func ? 2 : 1) * d.half;
    void *dB, *dT;
    if ((rc = ccfj_internal::ctx_scratch(ctx, 5, sizeof(float) * nf * Kp * d.nrhs, &dB))) return rc;
    const size_t tb = align256(sizeof(float) * (size_t)nwin * nf * nc);
    const size_t mb = align256(sizeof(float) * (size_t)nwin * nf);
    const size_t pb = align256(sizeof(float) * (size_t)fjk::mwfj_groups(nc) * nwin * nf);
    if ((rc = ccfj_internal::ctx_scratch(ctx, 6, tb + mb + pb + sizeof(float) * (size_t)nf, &dT)))
      return rc;
    d.B = (float*)dB;
    e = fjk::launch_mwfj_bpad(d, st);   // the coefficient kernels write the rest
    if (e != hipSuccess) return hip_fail(e, "padding kernel launch");
    d.T = (float*)dT;
    d.M = (float*)((char*)dT + tb);
    d.Mp = (float*)((char*)dT + tb + mb);
    d.bmax = (unsigned int*)((char*)dT + tb + mb + pb);
    e = fjk::launch_fill((float*)d.bmax, nf, 0.0f, st);   // bits 0: max |B| = 0
    if (e != hipSuccess) return hip_fail(e, "fill kernel launch");
  } else {
    // per-window staging: Re and Im spectra [nr][nf], one [nc][nf] result
    const size_t a = align256(sizeof(float) * (size_t)nr * nf);
    void* dst;
    if ((rc = ccfj_internal::ctx_scratch(ctx, 4, 2 * a + sizeof(float) * (size_t)nc * nf, &dst)))
      return rc;
    dre = (float*)dst;
    dim = (float*)((char*)dst + a);
    dtmp = (float*)((char*)dst + 2 * a);
  }

  // the spectra rows as the coefficient and staging kernels read them: the
  // Goertzel path stores the used bins only
  fjk::MwfjDev dd = d;
  int* rowbad = (int*)((char*)drr + o_bad);
  double2* tw = (double2*)((char*)drr + o_tw);
  if (direct) {
    dd.nbins = nf;
    dd.fstride = 1;
    dd.zhalf = 0;
    if ((e = fjk::launch_mwfj_rowbad(d, rowbad, st)) != hipSuccess)
      return hip_fail(e, "row check kernel");
    if ((e = fjk::launch_mwfj_twiddles(npts, tw, st)) != hipSuccess)
      return hip_fail(e, "twiddle kernel");
  }
  for (int w0 = 0; w0 < nwin; w0 += cw) {
    const int cwn = std::min(cw, nwin - w0);
    if (direct) {
      if ((e = fjk::launch_mwfj_goertzel(d, w0, cwn, rowbad, tw, (double2*)dspec, st)) !=
          hipSuccess)
        return hip_fail(e, "Goertzel kernel");
    } else if ((e = fjk::launch_window(d, w0, cwn, st)) != hipSuccess) {
      return hip_fail(e, "window kernel");
    }
    hipfftHandle plan = 0;
    if (direct) {
      // no transform
    } else if (d.zhalf) {
      if ((rc = ccfj_internal::cached_fft_plan(ctx, npts / 2, cwn * nr, HIPFFT_Z2Z, st, &plan)))
        return rc;
      if (hipfftExecZ2Z(plan, (hipfftDoubleComplex*)dxw, (hipfftDoubleComplex*)dspec,
                        HIPFFT_FORWARD) != HIPFFT_SUCCESS)
        return fail(CCFJ_EDEVICE, "hipfftExecZ2Z failed");
    } else {
      if ((rc = ccfj_internal::cached_fft_plan(ctx, npts, cwn * nr, HIPFFT_D2Z, st, &plan))) return rc;
      if (hipfftExecD2Z(plan, (hipfftDoubleReal*)dxw, (hipfftDoubleComplex*)dspec) !=
          HIPFFT_SUCCESS)
        return fail(CCFJ_EDEVICE, "hipfftExecD2Z failed");
    }
    if (itype == 0) {
      if ((e = fjk::launch_mwfj_coef(dd, w0, cwn, st)) != hipSuccess) return hip_fail(e, "coef kernel");
    } else {
      for (int wl = 0; wl < cwn; wl++) {
        if ((e = fjk::launch_mwfj_stage(dd, wl, dre, dim, st)) != hipSuccess)
          return hip_fail(e, "stage kernel");
        rc = ccfj_transform_dev(ctx, func ? CCFJ_MODE_ABS_H : CCFJ_MODE_ABS_J, 1, 1, dre, dim, d.r,
                                d.f, d.c, nr, nf, nc, dtmp, st);
        if (rc) return rc;
        e = fjk::launch_to_double(dtmp, (size_t)nc * nf, d.out + (size_t)(w0 + wl) * nc * nf, st);
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
  return CCFJ_OK;
}

}  // namespace

extern "C" int ccfj_mwfj_dev(ccfj_ctx* ctx, int itype, int func, int nr, int npts, int nwin,
                             int nf, int nc, int fstride, const double* u, const float* r,
                             const float* f, const float* c, const int* n1, const int* n2,
                             const int* delta, const int* xidx, const double* xrows, int nx,
                             double* out, void* stream) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  int rc = check_sizes(itype, func, nr, npts, nwin, nf, nc, fstride, nx);
  if (rc) return rc;
  if (!u || !r || !f || !c || !n1 || !n2 || !delta || !xidx || !out || (nx > 0 && !xrows))
    return fail(CCFJ_EINVAL, "null pointer");
  hipError_t e = hipSetDevice(ccfj_internal::ctx_device(ctx));
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = stream ? (hipStream_t)stream : ccfj_internal::ctx_stream(ctx);
  return mwfj_run(ctx, itype, func, nr, npts, nwin, nf, nc, fstride, u, r, f, c, n1, n2, delta,
                  xidx, xrows, nx, -1, out, st);
}

extern "C" int ccfj_mwfj(ccfj_ctx* ctx, int itype, int func, int nr, int npts, int nwin,
                         int nf, int nc, int fstride, const double* u, const float* r,
                         const float* f, const float* c, const int* n1, const int* n2,
                         const int* delta, const int* xidx, const double* xrows, int nx,
                         double* out) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  auto lk = ccfj_internal::ctx_lock(ctx);
  int rc = check_sizes(itype, func, nr, npts, nwin, nf, nc, fstride, nx);
  if (rc) return rc;
  if (!u || !r || !f || !c || !n1 || !n2 || !delta || !xidx || !out || (nx > 0 && !xrows))
    return fail(CCFJ_EINVAL, "null pointer");
  const int ndesc = nwin * nr;
  for (int k = 0; k < ndesc; k++)
    if (xidx[k] < -1 || xidx[k] >= nx)
      return fail(CCFJ_EINVAL, "xidx[%d] = %d outside [-1, %d)", k, xidx[k], nx);
  hipError_t e = hipSetDevice(ccfj_internal::ctx_device(ctx));
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = ccfj_internal::ctx_stream(ctx);

  // inputs staged in scratch slot 8 (one region, 256-byte aligned parts),
  // the result in slot 9
  const size_t o_u = 0;
  const size_t o_r = o_u + align256(sizeof(double) * (size_t)nr * npts);
  const size_t o_f = o_r + align256(sizeof(float) * nr);
  const size_t o_c = o_f + align256(sizeof(float) * nf);
  const size_t o_n1 = o_c + align256(sizeof(float) * nc);
  const size_t o_n2 = o_n1 + align256(sizeof(int) * ndesc);
  const size_t o_dl = o_n2 + align256(sizeof(int) * ndesc);
  const size_t o_xi = o_dl + align256(sizeof(int) * ndesc);
  const size_t o_xr = o_xi + align256(sizeof(int) * ndesc);
  const size_t total = o_xr + sizeof(double) * (size_t)nx * npts;
  void *din, *dout;
  if ((rc = ccfj_internal::ctx_scratch(ctx, 8, total, &din))) return rc;
  if ((rc = ccfj_internal::ctx_scratch(ctx, 9, sizeof(double) * (size_t)nwin * nc * nf, &dout)))
    return rc;
  char* b = (char*)din;
  struct Part { size_t off; const void* src; size_t bytes; };
  const Part parts[] = {{o_u, u, sizeof(double) * (size_t)nr * npts},
                        {o_r, r, sizeof(float) * nr},
                        {o_f, f, sizeof(float) * nf},
                        {o_c, c, sizeof(float) * nc},
                        {o_n1, n1, sizeof(int) * ndesc},
                        {o_n2, n2, sizeof(int) * ndesc},
                        {o_dl, delta, sizeof(int) * ndesc},
                        {o_xi, xidx, sizeof(int) * ndesc},
                        {o_xr, xrows, sizeof(double) * (size_t)nx * npts}};
  for (const Part& q : parts) {
    if (!q.bytes) continue;
    if ((rc = ccfj_internal::copy_h2d(ctx, b + q.off, q.src, q.bytes, st))) return rc;
  }
  rc = mwfj_run(ctx, itype, func, nr, npts, nwin, nf, nc, fstride, (const double*)(b + o_u),
                (const float*)(b + o_r), (const float*)(b + o_f), (const float*)(b + o_c),
                (const int*)(b + o_n1), (const int*)(b + o_n2), (const int*)(b + o_dl),
                (const int*)(b + o_xi), nx > 0 ? (const double*)(b + o_xr) : nullptr, nx,
                support_sum(nr, npts, nwin, n1, n2, xidx, nx), (double*)dout, st);
  if (rc) {
    (void)hipStreamSynchronize(st);
    return rc;
  }
  return ccfj_internal::copy_d2h(ctx, out, dout, sizeof(double) * (size_t)nwin * nc * nf, st);
}
