// fj_cc.hip -- ambient-noise cross-correlation on the device: the producer
// of the spectra fj_noise consumes.
//
// What is computed (reference: src/CrossCorr.cpp:33-159): the record of
// every station is cut into windows of fftlen samples every steplen
// samples; each window's real FFT is optionally whitened (divided by
// max(|X|, 1e-8)); for every station pair (a, b) the products
// conj(X_a) X_b at bins j*fstride (j < nf) are summed over the windows in
// which both stations have data, then divided by that count.
//
// Layout (device):
//   data  f32 [nsta][npts]            (one-bit: replaced by its sign first)
//   seg   f32 [cw][nsta][fftlen]      windows of a chunk, gathered
//   spec  c64 [cw][nsta][fftlen/2+1]  their rFFT (hipFFT, single precision)
//   S     c64 [cw][nsta][nf]          bins j*fstride, whitened
//   acc   c64 [npairs][nf]            running sums over windows
//   use   u8  [nshifts][nsta]         station has the whole window
//   count f32 [npairs]                windows both stations of a pair hold
//
// The pair kernel reads S rows of two stations per pair; a chunk's S
// (cw * nsta * nf * 8 bytes) stays in L2/MALL while all pairs sweep it, and
// acc is read and written once per chunk of cw windows, not per window.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fj_cc.h"

namespace fjk {

__global__ void k_cc_onebit(float* __restrict__ x, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = x[i];
  x[i] = v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f);
}

// seg[wl][i][t] = data[i][(w0 + wl) * steplen + t]; grid (t blocks, i, wl)
__global__ void k_cc_gather(const float* __restrict__ data, int npts, int fftlen, int steplen,
                            int w0, float* __restrict__ seg) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  const int wl = blockIdx.z;
  if (t >= fftlen) return;
  const size_t src = (size_t)i * npts + (size_t)(w0 + wl) * steplen + t;
  seg[((size_t)wl * gridDim.y + i) * fftlen + t] = data[src];
}

// S[row][j] = spec[row][j * fstride] (whitened); rows = cw * nsta
__global__ void k_cc_spec(const float2* __restrict__ spec, int nbins, int nf, int fstride,
                          int whiten, float2* __restrict__ S) {
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  const size_t row = blockIdx.x;
  if (j >= nf) return;
  float2 v = spec[row * nbins + (size_t)j * fstride];
  if (whiten) {
    // as the reference: |X| in float, floored at 1e-8, then a division
    const float a = fmaxf(sqrtf(v.x * v.x + v.y * v.y), 1e-8f);
    v.x = v.x / a;
    v.y = v.y / a;
  }
  S[row * nf + j] = v;
}

// use[k][i] = station i holds window k: k*steplen >= start_i and
// k*steplen + fftlen < end_i (src/CrossCorr.cpp:91-98)
__global__ void k_cc_use(const int* __restrict__ startend, int nsta, size_t n, int steplen,
                         int fftlen, unsigned char* __restrict__ use) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int i = (int)(e % nsta);
  const long long t0 = (long long)(e / nsta) * steplen;
  use[e] = t0 >= startend[2 * i] && t0 + fftlen < startend[2 * i + 1];
}

// count[p] = windows that both stations of pair p hold (0 for a pair that
// names a station outside [0, nsta))
__global__ void k_cc_count(const unsigned char* __restrict__ use, const int* __restrict__ pairs,
                           int npairs, int nsta, int nshifts, float* __restrict__ count) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int a = pairs[2 * p], b = pairs[2 * p + 1];
  int n = 0;
  if (a >= 0 && a < nsta && b >= 0 && b < nsta)
    for (int k = 0; k < nshifts; k++) n += use[(size_t)k * nsta + a] & use[(size_t)k * nsta + b];
  count[p] = (float)n;   // the reference divides the float sums by it
}

// acc[p][j] += sum over wl < cw with use[w0+wl][a] && use[w0+wl][b] of
// conj(S[wl][a][j]) S[wl][b][j].  Block: one pair (grid x), 256 bins (y).
// A pair naming a station outside [0, nsta) adds nothing.
__global__ __launch_bounds__(256) void k_cc_pairs(const float2* __restrict__ S,
                                                  const unsigned char* __restrict__ use,
                                                  const int* __restrict__ pairs, int nsta,
                                                  int nf, int w0, int cw,
                                                  float2* __restrict__ acc) {
  const int p = blockIdx.x;
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  const int a = pairs[2 * p], b = pairs[2 * p + 1];
  if (j >= nf || a < 0 || a >= nsta || b < 0 || b >= nsta) return;
  float re = 0.0f, im = 0.0f;
  for (int wl = 0; wl < cw; wl++) {
    const unsigned char* u = use + (size_t)(w0 + wl) * nsta;
    if (!(u[a] && u[b])) continue;   // uniform over the block
    const float2 xa = S[((size_t)wl * nsta + a) * nf + j];
    const float2 xb = S[((size_t)wl * nsta + b) * nf + j];
    re = __builtin_fmaf(xa.x, xb.x, __builtin_fmaf(xa.y, xb.y, re));
    im = __builtin_fmaf(xa.x, xb.y, __builtin_fmaf(-xa.y, xb.x, im));
  }
  float2* o = acc + (size_t)p * nf + j;
  float2 v = *o;
  v.x += re;
  v.y += im;
  *o = v;
}

// acc[k] = (re[k], im[k]): running sums from separate arrays
__global__ void k_cc_interleave(const float* __restrict__ re, const float* __restrict__ im,
                                size_t n, float2* __restrict__ acc) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) acc[k] = make_float2(re[k], im[k]);
}

// ncfsr/ncfsi[p][j] = acc[p][j] / count[p] where count > 0, else acc
// (the untouched running sums, as the reference leaves them)
__global__ void k_cc_finish(const float2* __restrict__ acc, const float* __restrict__ count,
                            int nf, float* __restrict__ re, float* __restrict__ im) {
  const int p = blockIdx.x;
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= nf) return;
  const float n = count[p];
  const float2 v = acc[(size_t)p * nf + j];
  re[(size_t)p * nf + j] = n > 0.0f ? v.x / n : v.x;
  im[(size_t)p * nf + j] = n > 0.0f ? v.y / n : v.y;
}

hipError_t launch_cc_onebit(float* x, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cc_onebit, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n);
  return hipGetLastError();
}

hipError_t launch_cc_gather(const float* data, int nsta, int npts, int fftlen, int steplen, int w0,
                            int cw, float* seg, hipStream_t st) {
  hipLaunchKernelGGL(k_cc_gather, dim3((fftlen + 255) / 256, nsta, cw), dim3(256), 0, st, data,
                     npts, fftlen, steplen, w0, seg);
  return hipGetLastError();
}

hipError_t launch_cc_spec(const void* spec, int rows, int nbins, int nf, int fstride, int whiten,
                          void* S, hipStream_t st) {
  hipLaunchKernelGGL(k_cc_spec, dim3(rows, (nf + 255) / 256), dim3(256), 0, st,
                     (const float2*)spec, nbins, nf, fstride, whiten, (float2*)S);
  return hipGetLastError();
}

hipError_t launch_cc_use(const int* startend, int nsta, int nshifts, int steplen, int fftlen,
                         unsigned char* use, hipStream_t st) {
  const size_t n = (size_t)nshifts * nsta;
  hipLaunchKernelGGL(k_cc_use, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, startend, nsta,
                     n, steplen, fftlen, use);
  return hipGetLastError();
}

hipError_t launch_cc_count(const unsigned char* use, const int* pairs, int npairs, int nsta,
                           int nshifts, float* count, hipStream_t st) {
  hipLaunchKernelGGL(k_cc_count, dim3((npairs + 255) / 256), dim3(256), 0, st, use, pairs, npairs,
                     nsta, nshifts, count);
  return hipGetLastError();
}

hipError_t launch_cc_pairs(const void* S, const unsigned char* use, const int* pairs, int npairs,
                           int nsta, int nf, int w0, int cw, void* acc, hipStream_t st) {
  hipLaunchKernelGGL(k_cc_pairs, dim3(npairs, (nf + 255) / 256), dim3(256), 0, st,
                     (const float2*)S, use, pairs, nsta, nf, w0, cw, (float2*)acc);
  return hipGetLastError();
}

hipError_t launch_cc_interleave(const float* re, const float* im, size_t n, void* acc,
                                hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cc_interleave, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, re, im,
                     n, (float2*)acc);
  return hipGetLastError();
}

hipError_t launch_cc_finish(const void* acc, const float* count, int npairs, int nf, float* re,
                            float* im, hipStream_t st) {
  hipLaunchKernelGGL(k_cc_finish, dim3(npairs, (nf + 255) / 256), dim3(256), 0, st,
                     (const float2*)acc, count, nf, re, im);
  return hipGetLastError();
}

}  // namespace fjk
