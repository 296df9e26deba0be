// fj_plain64.hip -- the plain single-instance FP64 GPU F-J (speed-up baseline).
//
// Structure of the reference GPU engine (src/FJgpu.cu:160-313, hosts
// :390-483): one thread per (f,c) in 32x16 blocks, a serial loop over the
// station segments reading U_f[i + ir*nf], the special functions evaluated
// at both segment ends, and device buffers allocated, copied and freed on
// every call.  Unlike the reference it is FP64 throughout (k, k*r, the
// accumulator), uses the ROCm device library's double j0/j1/y0/y1, and the
// J-basis linear-segment kernel uses the true int J0 (the reference's
// float integral_c truncates it, src/FJgpu.cu:345-354).  This is the
// "plain FP64 GPU F-J" denominator of the speed-up, not the product path.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/ccfj.h"

namespace {

constexpr double kPi = 3.141592653589793;

// Struve H0, H1 by the Zhang & Jin (1996) method: power series up to x = 20,
// asymptotic series of H - Y plus their own Y polynomials beyond.
__device__ double struve_h0(double x) {
  double s = 1.0, r = 1.0;
  if (x <= 20.0) {
    for (int k = 1; k < 61; k++) {
      r = -r * x / (2.0 * k + 1.0) * x / (2.0 * k + 1.0);
      s += r;
      if (fabs(r) < fabs(s) * 1.0e-12) break;
    }
    return 2.0 * x / kPi * s;
  }
  int km = x >= 50.0 ? 25 : (int)(0.5 * (x + 1.0));
  for (int k = 1; k <= km; k++) {
    r = -r * (2.0 * k - 1.0) * (2.0 * k - 1.0) / x / x;
    s += r;
    if (fabs(r) < fabs(s) * 1.0e-12) break;
  }
  double t = 4.0 / x, t2 = t * t;
  double p0 = ((((-.37043e-5 * t2 + .173565e-4) * t2 - .487613e-4) * t2 + .17343e-3) * t2 -
               0.1753062e-2) * t2 + .3989422793;
  double q0 = t * (((((.32312e-5 * t2 - 0.142078e-4) * t2 + 0.342468e-4) * t2 - 0.869791e-4) *
                        t2 + 0.4564324e-3) * t2 - 0.0124669441);
  double ta0 = x - 0.25 * kPi;
  return 2.0 / (kPi * x) * s + 2.0 / sqrt(x) * (p0 * sin(ta0) + q0 * cos(ta0));
}

__device__ double struve_h1(double x) {
  double r = 1.0, s;
  if (x <= 20.0) {
    s = 0.0;
    for (int k = 1; k <= 60; k++) {
      r = -r * x * x / (4.0 * k * k - 1.0);
      s += r;
      if (fabs(r) < fabs(s) * 1.0e-12) break;
    }
    return -2.0 / kPi * s;
  }
  s = 1.0;
  int km = x > 50.0 ? 25 : (int)(0.5 * x);
  for (int k = 1; k <= km; k++) {
    r = -r * (4.0 * k * k - 1.0) / (x * x);
    s += r;
    if (fabs(r) < fabs(s) * 1.0e-12) break;
  }
  double t = 4.0 / x, t2 = t * t;
  double p1 = ((((0.42414e-5 * t2 - 0.20092e-4) * t2 + 0.580759e-4) * t2 - 0.223203e-3) * t2 +
               0.29218256e-2) * t2 + 0.3989422819;
  double q1 = t * (((((-0.36594e-5 * t2 + 0.1622e-4) * t2 - 0.398708e-4) * t2 + 0.1064741e-3) *
                        t2 - 0.63904e-3) * t2 + 0.0374008364);
  double ta1 = x - 0.75 * kPi;
  return 2.0 / kPi * (1.0 + s / (x * x)) + 2.0 / sqrt(x) * (p1 * sin(ta1) + q1 * cos(ta1));
}

template <bool Y>
__device__ inline double z0(double x) { return Y ? y0(x) : j0(x); }
template <bool Y>
__device__ inline double z1(double x) { return Y ? y1(x) : j1(x); }

template <bool Y, int ITYPE>
__global__ void k_plain(const float* __restrict__ U, const float* __restrict__ r,
                        const float* __restrict__ f, const float* __restrict__ c, int nr,
                        int nf, int nc, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= nf || j >= nc) return;
  const double k = 2.0 * kPi * (double)f[i] / (double)c[j];
  double acc = 0.0;
  for (int ir = 1; ir < nr; ir++) {
    const double g1 = U[i + (size_t)(ir - 1) * nf], g2 = U[i + (size_t)ir * nf];
    const double r1 = r[ir - 1], r2 = r[ir];
    const double dr0 = fmax(r2 - r1, 0.1);
    const double x1 = k * r1, x2 = k * r2;
    if (ITYPE == 0) {
      acc += (g1 * z0<Y>(x1) * r1 + g2 * z0<Y>(x2) * r2) * dr0 * 0.5;
    } else {
      const double b = (g2 - g1) / dr0;
      const double a = g1 - r1 * b;
      const double z01 = z0<Y>(x1), z02 = z0<Y>(x2), z11 = z1<Y>(x1), z12 = z1<Y>(x2);
      acc += a * (r2 * z12 - r1 * z11) / k;
      acc += b * (r2 * r2 * z12 - r1 * r1 * z11) / k;
      acc += b * (r2 * z02 - r1 * z01) / (k * k);
      const double B2 = x2 * z02 + kPi * x2 * (z12 * struve_h0(x2) - z02 * struve_h1(x2)) / 2;
      const double B1 = x1 * z01 + kPi * x1 * (z11 * struve_h0(x1) - z01 * struve_h1(x1)) / 2;
      acc -= b * (B2 - B1) / (k * k * k);
    }
  }
  out[i + (size_t)j * nf] = (float)acc;
}

}  // namespace

extern "C" int ccfj_plain_fp64(int device, int ybasis, int itype, const float* U,
                               const float* r, const float* f, const float* c, int nr,
                               int nf, int nc, float* out, double* kernel_ms) {
  if (itype != 0 && itype != 1) return CCFJ_EINVAL;
  if (nr < 0 || nf <= 0 || nc <= 0) return CCFJ_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return CCFJ_EDEVICE;
  if (device >= 0 && hipSetDevice(device) != hipSuccess) return CCFJ_EDEVICE;
  float *dU = nullptr, *dr = nullptr, *df = nullptr, *dc = nullptr, *dout = nullptr;
  const size_t nu = (size_t)nr * nf;
  int rc = CCFJ_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipMalloc(&dU, sizeof(float) * (nu ? nu : 1)) != hipSuccess ||
      hipMalloc(&dr, sizeof(float) * (nr ? nr : 1)) != hipSuccess ||
      hipMalloc(&df, sizeof(float) * nf) != hipSuccess ||
      hipMalloc(&dc, sizeof(float) * nc) != hipSuccess ||
      hipMalloc(&dout, sizeof(float) * (size_t)nf * nc) != hipSuccess) {
    rc = CCFJ_ENOMEM;
  } else {
    hipError_t e = hipMemcpy(dU, U, sizeof(float) * nu, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dr, r, sizeof(float) * nr, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(df, f, sizeof(float) * nf, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dc, c, sizeof(float) * nc, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) {
      rc = CCFJ_EDEVICE;
    } else {
      dim3 block(32, 16), grid((nf + 31) / 32, (nc + 15) / 16);
      (void)hipEventRecord(e0, 0);
      if (ybasis) {
        if (itype) hipLaunchKernelGGL((k_plain<true, 1>), grid, block, 0, 0, dU, dr, df, dc, nr, nf, nc, dout);
        else hipLaunchKernelGGL((k_plain<true, 0>), grid, block, 0, 0, dU, dr, df, dc, nr, nf, nc, dout);
      } else {
        if (itype) hipLaunchKernelGGL((k_plain<false, 1>), grid, block, 0, 0, dU, dr, df, dc, nr, nf, nc, dout);
        else hipLaunchKernelGGL((k_plain<false, 0>), grid, block, 0, 0, dU, dr, df, dc, nr, nf, nc, dout);
      }
      (void)hipEventRecord(e1, 0);
      if (hipGetLastError() != hipSuccess) rc = CCFJ_EDEVICE;
      if (hipMemcpy(out, dout, sizeof(float) * (size_t)nf * nc, hipMemcpyDeviceToHost) != hipSuccess)
        rc = CCFJ_EDEVICE;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (kernel_ms) *kernel_ms = ms;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(dU);
  (void)hipFree(dr);
  (void)hipFree(df);
  (void)hipFree(dc);
  (void)hipFree(dout);
  return rc;
}
