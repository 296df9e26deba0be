// fj_hilbert.hip -- Hilbert transform along frequency and fj_noise on the
// device.
//
// The reference transforms every row of the spectra with
// scipy.fftpack.hilbert in a Python loop (src/ccfj.pyx:217-219):
// y = ifft(i sign(k) fft(x)), zero and (even length) Nyquist coefficients
// dropped, computed in double.  For a real row that is
// y = irfft(Y), Y_k = i X_k for 0 < k < n/2, Y_0 = Y_{n/2} = 0, X = rfft(x);
// here: double rFFT (hipFFT D2Z), one multiply kernel, inverse (Z2D),
// scale and cast to float32 -- the float32 the reference stores into ufi.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/ccfj.h"
#include "ccfj_internal.h"

using ccfj_internal::fail;
using ccfj_internal::hip_fail;

namespace {

__global__ void k_to_double_rows(const float* __restrict__ x, int64_t n, double* __restrict__ y) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[k] = (double)x[k];
}

// Y_k = i X_k (0 < k < n/2), zero at k = 0 and, for even n, k = n/2
__global__ void k_times_i(double2* __restrict__ X, int rows, int n) {
  const int nb = n / 2 + 1;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)rows * nb) return;
  const int bin = (int)(k % nb);
  double2 z = X[k];
  if (bin == 0 || (n % 2 == 0 && bin == n / 2)) {
    z.x = 0.0;
    z.y = 0.0;
  } else {
    const double re = z.x;
    z.x = -z.y;
    z.y = re;
  }
  X[k] = z;
}

__global__ void k_scale_to_float(const double* __restrict__ y, int64_t n, double s,
                                 float* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = (float)(y[k] * s);
}

// out[s][i] = in[s][i * stride], i < nf   (the reference's uf[:, ::fstride])
__global__ void k_stride_cols(const float* __restrict__ in, int ncol, int stride, int nf,
                              float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= nf) return;
  out[(size_t)s * nf + i] = in[(size_t)s * ncol + (size_t)i * stride];
}

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

}  // namespace


namespace {

int get_plan(ccfj_ctx* ctx, int n, int batch, bool forward, hipStream_t st, hipfftHandle* h) {
  return ccfj_internal::cached_fft_plan(ctx, n, batch, forward ? HIPFFT_D2Z : HIPFFT_Z2D, st, h);
}

// Hilbert transform of rows x -> y (device, float32), double inside; the
// double work arrays are the context's scratch buffers.
int hilbert_rows(ccfj_ctx* ctx, const float* x, float* y, int rows, int n, hipStream_t st) {
  if (rows <= 0 || n <= 0) return CCFJ_OK;
  const int nb = n / 2 + 1;
  const size_t per_row = (size_t)n * 8 + (size_t)nb * 16;
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>(rows, ((size_t)1 << 31) / per_row));
  void *dr, *dc;
  int rc = ccfj_internal::ctx_scratch(ctx, 0, (size_t)chunk * n * 8, &dr);
  if (!rc) rc = ccfj_internal::ctx_scratch(ctx, 1, (size_t)chunk * nb * 16, &dc);
  if (rc) return rc;
  for (int r0 = 0; r0 < rows; r0 += chunk) {
    const int m = std::min(chunk, rows - r0);
    hipfftHandle fwd, inv;
    if ((rc = get_plan(ctx, n, m, true, st, &fwd)) || (rc = get_plan(ctx, n, m, false, st, &inv)))
      return rc;
    const int64_t nr_ = (int64_t)m * n, nc_ = (int64_t)m * nb;
    hipLaunchKernelGGL(k_to_double_rows, dim3((unsigned)((nr_ + 255) / 256)), dim3(256), 0, st,
                       x + (size_t)r0 * n, nr_, (double*)dr);
    bool ok = hipGetLastError() == hipSuccess;
    ok = ok && hipfftExecD2Z(fwd, (hipfftDoubleReal*)dr, (hipfftDoubleComplex*)dc) == HIPFFT_SUCCESS;
    if (ok) {
      hipLaunchKernelGGL(k_times_i, dim3((unsigned)((nc_ + 255) / 256)), dim3(256), 0, st,
                         (double2*)dc, m, n);
      ok = hipGetLastError() == hipSuccess;
    }
    ok = ok && hipfftExecZ2D(inv, (hipfftDoubleComplex*)dc, (hipfftDoubleReal*)dr) == HIPFFT_SUCCESS;
    if (ok) {
      hipLaunchKernelGGL(k_scale_to_float, dim3((unsigned)((nr_ + 255) / 256)), dim3(256), 0, st,
                         (const double*)dr, nr_, 1.0 / n, y + (size_t)r0 * n);
      ok = hipGetLastError() == hipSuccess;
    }
    if (!ok) return fail(CCFJ_EDEVICE, "Hilbert transform launch failed");
  }
  return CCFJ_OK;
}

}  // namespace

extern "C" int ccfj_hilbert_dev(ccfj_ctx* ctx, const float* x, float* y, int rows, int n,
                                void* stream) {
  if (!ctx || (!x && rows > 0) || (!y && rows > 0)) return fail(CCFJ_EINVAL, "null argument");
  hipError_t e = hipSetDevice(ccfj_internal::ctx_device(ctx));
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = stream ? (hipStream_t)stream : ccfj_internal::ctx_stream(ctx);
  return hilbert_rows(ctx, x, y, rows, n, st);
}

extern "C" int ccfj_fj_noise(ccfj_ctx* ctx, int itype, int func, int fstride, const float* uf,
                             const float* r, const float* f, const float* c, int nr, int ncol,
                             int nf, int nc, float* out) {
  if (!ctx) return fail(CCFJ_EINVAL, "null context");
  auto lk = ccfj_internal::ctx_lock(ctx);
  if (func != 0 && func != 1) return fail(CCFJ_EINVAL, "func must be 0 or 1, got %d", func);
  if (itype != 0 && itype != 1) return fail(CCFJ_EINVAL, "itype must be 0 or 1, got %d", itype);
  if (fstride < 1 || nr < 0 || ncol < 0 || nf < 0 || nc < 0)
    return fail(CCFJ_EINVAL, "bad sizes");
  if ((int64_t)(nf - 1) * fstride >= ncol && nf > 0 && nr > 0)
    return fail(CCFJ_EINVAL, "nf=%d columns with stride %d exceed the %d input columns", nf, fstride, ncol);
  if (nf == 0 || nc == 0) return CCFJ_OK;
  hipError_t e = hipSetDevice(ccfj_internal::ctx_device(ctx));
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t st = ccfj_internal::ctx_stream(ctx);
  Bufs b;
  void *du = nullptr, *dh = nullptr, *dU = nullptr, *dH = nullptr, *dr, *df, *dc, *dout;
  int rc = CCFJ_OK;
  const size_t nfull = (size_t)nr * ncol, nst = (size_t)nr * nf;
  if (!rc) rc = b.get(sizeof(float) * (nfull + 1), &du);
  if (!rc) rc = b.get(sizeof(float) * (nst + 1), &dU);
  if (!rc && func == 1) rc = b.get(sizeof(float) * (nfull + 1), &dh);
  if (!rc && func == 1) rc = b.get(sizeof(float) * (nst + 1), &dH);
  if (!rc) rc = b.get(sizeof(float) * (nr + 1), &dr);
  if (!rc) rc = b.get(sizeof(float) * nf, &df);
  if (!rc) rc = b.get(sizeof(float) * nc, &dc);
  if (!rc) rc = b.get(sizeof(float) * (size_t)nf * nc, &dout);
  if (rc) return rc;
  if ((rc = ccfj_internal::copy_h2d(ctx, du, uf, sizeof(float) * nfull, st))) return rc;
  e = hipMemcpyAsync(dr, r, sizeof(float) * nr, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(df, f, sizeof(float) * nf, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dc, c, sizeof(float) * nc, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync H2D");
  if (nr > 0) {
    dim3 grid((nf + 255) / 256, nr);
    hipLaunchKernelGGL(k_stride_cols, grid, dim3(256), 0, st, (const float*)du, ncol, fstride, nf,
                       (float*)dU);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "stride kernel");
    if (func == 1) {
      if ((rc = hilbert_rows(ctx, (const float*)du, (float*)dh, nr, ncol, st))) return rc;
      hipLaunchKernelGGL(k_stride_cols, grid, dim3(256), 0, st, (const float*)dh, ncol, fstride,
                         nf, (float*)dH);
      if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "stride kernel");
    }
  }
  rc = ccfj_transform_dev(ctx, func ? CCFJ_MODE_J_MINUS_Y : CCFJ_MODE_J, itype, 1,
                          (const float*)dU, (const float*)dH, (const float*)dr, (const float*)df,
                          (const float*)dc, nr, nf, nc, (float*)dout, st);
  if (rc) return rc;
  return ccfj_internal::copy_d2h(ctx, out, dout, sizeof(float) * (size_t)nf * nc, st);
}
