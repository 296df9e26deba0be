// fj_synth.hip -- synthetic multimode noise fields and their ridge labels,
// generated on the device for training the band-proposal CNN, and the
// item-list F-J kernel that evaluates only proposed velocity bands.
//
// Field of sample b (nr stations, distances r_b sorted, shared f grid):
//   U_b[s][i] = sum_m A_m J0(2 pi f_i r_s / c_m(f_i)) + sigma_b n(b, s, i)
//   c_m(f) = cinf_m + d_m exp(-f / phi_m)   (monotone decreasing)
// n is a standard normal from a counter-based hash (splitmix64 + Box-Muller)
// of (seed, b, s, i), so a field is reproducible from its seed alone.
// Label of sample b on the coarse grid (velocity h, frequency i):
//   1 if |c_h - c_m(f_i)| <= halfwidth for a mode with A_m >= amin, else 0.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fj_lane.h"
#include "fj_synth.h"

namespace fjk {

using namespace fjl;

__device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ inline float gauss(uint64_t key) {
  const uint64_t a = splitmix64(key), b = splitmix64(key ^ 0xD1B54A32D192ED03ull);
  const double u1 = ((a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
  return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

// modes: [B][kMaxModes][4] = (cinf, d, phi, amp); amp = 0 for unused modes
__global__ void k_synth_fields(const float* __restrict__ r, const float* __restrict__ f,
                               const float* __restrict__ modes, const float* __restrict__ sigma,
                               int nr, int nf, uint64_t seed, float* __restrict__ U) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  const int b = blockIdx.z;
  if (i >= nf) return;
  const double fi = f[i];
  const double rs = r[(size_t)b * nr + s];
  double acc = 0.0;
  for (int m = 0; m < kMaxModes; m++) {
    const float* p = modes + ((size_t)b * kMaxModes + m) * 4;
    if (p[3] == 0.0f) continue;
    const double cm = p[0] + p[1] * exp(-fi / p[2]);
    const double q = fi / cm;                 // cycles per metre
    const double xd = q * kTwoPi * rs;
    const float x = (float)xd;
    float J0 = 0.0f, Y0;
    if (x >= kSplit) {
      bessel0_large<true, false>(x, reduce_phase(q, rs), &J0, &Y0);
    } else {
      double y, L;
      J0 = bessel_small<false, false, FIT_TABLE>(xd, &y, &L).j0;
    }
    acc += p[3] * J0;
  }
  const uint64_t key = seed ^ (((uint64_t)b * 1000003ull + s) * 1000033ull + i);
  U[((size_t)b * nr + s) * nf + i] = (float)acc + sigma[b] * gauss(key);
}

__global__ void k_synth_labels(const float* __restrict__ cgrid, const float* __restrict__ f,
                               const float* __restrict__ modes, int H, int nf, float halfwidth,
                               float amin, float* __restrict__ Y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = blockIdx.y;
  const int b = blockIdx.z;
  if (i >= nf) return;
  const float c = cgrid[h];
  float lab = 0.0f;
  for (int m = 0; m < kMaxModes; m++) {
    const float* p = modes + ((size_t)b * kMaxModes + m) * 4;
    if (p[3] < amin) continue;
    const float cm = p[0] + p[1] * expf(-f[i] / p[2]);
    if (fabsf(c - cm) <= halfwidth) lab = 1.0f;
  }
  Y[((size_t)b * H + h) * nf + i] = lab;
}

// out[b][h][w] = max over s < S of |in[b][h*S + s][w]|: the coarse spectrum
// as the envelope of S sub-velocities per cell (a ridge narrower than a
// cell still shows).
__global__ void k_pool_abs_max(const float* __restrict__ in, int H, int S, int W,
                               float* __restrict__ out) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = blockIdx.y;
  const int b = blockIdx.z;
  if (w >= W) return;
  const float* src = in + ((size_t)b * H * S + (size_t)h * S) * W + w;
  float m = fabsf(src[0]);
  for (int s = 1; s < S; s++) m = fmaxf(m, fabsf(src[(size_t)s * W]));
  out[((size_t)b * H + h) * W + w] = m;
}

// out[b][h][w] = max over s in [lo_h, hi_h) of |in[b][s][w]|, in [B][N][W],
// ranges[h] = (lo_h, hi_h): the coarse spectrum as the envelope of the
// sub-velocities that fall in each cell (any number per cell).  Ranges are
// clipped to [0, N); an empty one gives 0.
__global__ void k_pool_abs_max_ranges(const float* __restrict__ in, int N, int H, int W,
                                      const int* __restrict__ ranges, float* __restrict__ out) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = blockIdx.y;
  const int b = blockIdx.z;
  if (w >= W) return;
  int lo = ranges[2 * h], hi = ranges[2 * h + 1];
  lo = lo < 0 ? 0 : lo;
  hi = hi > N ? N : hi;
  const float* src = in + (size_t)b * N * W + w;
  float m = 0.0f;
  for (int s = lo; s < hi; s++) m = fmaxf(m, fabsf(src[(size_t)s * W]));
  out[((size_t)b * H + h) * W + w] = m;
}

// Y[b][h*][w] = 1 at the velocity cell h* holding the column maximum of X[b]
// (first on ties); other entries unchanged.
__global__ void k_label_argmax(const float* __restrict__ X, int B, int H, int W,
                               float* __restrict__ Y) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (w >= W || b >= B) return;
  const float* x = X + (size_t)b * H * W + w;
  int best = 0;
  float m = x[0];
  for (int h = 1; h < H; h++) {
    const float v = x[(size_t)h * W];
    if (v > m) {
      m = v;
      best = h;
    }
  }
  Y[((size_t)b * H + best) * W + w] = 1.0f;
}

// F-J on a list of (frequency, first velocity, count <= 256) items; lanes
// beyond count idle.  out[c][f] is written only for listed entries; an item
// outside the grid is skipped (items live in device memory, so the host
// cannot check them).  Fewer than two stations: zeros, as the full grid.
template <int MODE, int ITYPE>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(FjOcc<MODE, ITYPE, true>::waves, 8))) void k_fj_items(
    Stations S, const float* __restrict__ coef,
                                                 const float* __restrict__ fv,
                                                 const float* __restrict__ cv,
                                                 const int* __restrict__ items, int nr, int nrp,
                                                 int nf, int nc, float* __restrict__ out) {
  const int* it = items + 3 * (size_t)blockIdx.x;
  const int fi = it[0], c0 = it[1];
  const int cnt = min(it[2], nc - c0);   // item clipped to the grid
  // uniform over the workgroup: a skipped item exits as a whole
  if (fi < 0 || fi >= nf || c0 < 0 || cnt <= 0) return;
  if ((int)(threadIdx.x & ~63u) >= cnt) return;   // whole waves past the item
  const int j = c0 + min((int)threadIdx.x, cnt - 1);   // surplus lanes repeat the last
  const bool store = (int)threadIdx.x < cnt;
  if (nr < 2) {
    if (store) out[(size_t)j * nf + fi] = 0.0f;
    return;
  }
  const double q = (double)fv[fi] / (double)cv[j];
  const float* __restrict__ cf = coef + (size_t)fi * coef_block(nrp, MODE, ITYPE);
  const float v = ITYPE == 0 ? trapz_lane<MODE>(S, cf, nrp, q) : linseg_lane<MODE, true>(S, cf, nrp, q);
  if (store) out[(size_t)j * nf + fi] = v;
}

hipError_t launch_synth_fields(const float* r, const float* f, const float* modes,
                               const float* sigma, int B, int nr, int nf, uint64_t seed, float* U,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_synth_fields, dim3((nf + 255) / 256, nr, B), dim3(256), 0, st, r, f, modes,
                     sigma, nr, nf, seed, U);
  return hipGetLastError();
}

hipError_t launch_synth_labels(const float* cgrid, const float* f, const float* modes, int B,
                               int H, int nf, float halfwidth, float amin, float* Y,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_synth_labels, dim3((nf + 255) / 256, H, B), dim3(256), 0, st, cgrid, f,
                     modes, H, nf, halfwidth, amin, Y);
  return hipGetLastError();
}

hipError_t launch_pool_abs_max_ranges(const float* in, int B, int N, int H, int W,
                                      const int* ranges, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_pool_abs_max_ranges, dim3((W + 255) / 256, H, B), dim3(256), 0, st, in, N,
                     H, W, ranges, out);
  return hipGetLastError();
}

hipError_t launch_pool_abs_max(const float* in, int B, int H, int S, int W, float* out,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_pool_abs_max, dim3((W + 255) / 256, H, B), dim3(256), 0, st, in, H, S, W,
                     out);
  return hipGetLastError();
}

hipError_t launch_label_argmax(const float* X, int B, int H, int W, float* Y, hipStream_t st) {
  hipLaunchKernelGGL(k_label_argmax, dim3((W + 255) / 256, B), dim3(256), 0, st, X, B, H, W, Y);
  return hipGetLastError();
}

hipError_t launch_fj_items(int mode, int itype, const Stations& S, const float* coef,
                           const float* f, const float* c, const int* items, int nitems, int nr,
                           int nrp, int nf, int nc, float* out, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
#define FJ_ITEMS(M, I)                                                                         \
  hipLaunchKernelGGL((k_fj_items<M, I>), dim3(nitems), dim3(256), 0, st, S, coef, f, c,        \
                     items, nr, nrp, nf, nc, out)
  if (mode == MODE_J && itype == 0) FJ_ITEMS(MODE_J, 0);
  else if (mode == MODE_J && itype == 1) FJ_ITEMS(MODE_J, 1);
  else if (mode == MODE_J_MINUS_Y && itype == 0) FJ_ITEMS(MODE_J_MINUS_Y, 0);
  else if (mode == MODE_J_MINUS_Y && itype == 1) FJ_ITEMS(MODE_J_MINUS_Y, 1);
  else return hipErrorInvalidValue;
#undef FJ_ITEMS
  return hipGetLastError();
}

}  // namespace fjk
