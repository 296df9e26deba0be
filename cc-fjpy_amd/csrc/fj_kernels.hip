// fj_kernels.hip -- F-J transform kernels for gfx950 (MI355X).
//
// The per-(f,c) arithmetic lives in fj_lane.h (shared with the host build);
// this file maps it onto the GPU.
//
// Data layout in HBM (prepared on the device by k_stations and k_prep;
// nrp = nr rounded up to the station block of 8, padding = last distance
// with zero coefficients and zero gaps):
//   rr[nrp + 8]       double   station distances, ascending
//   hh[nrp + 8]       float    r_s - r_{s-1} (hh[0] = 0)
//   ir, isr[nrp + 8]  float    1 / r_s, 1 / sqrt(r_s)
//   bw[nrp + 8]       uint32   block words of the far / smooth / anchor
//                              tests (fjl::station_block_word)
//   dr[nrp + 8]       float    r_s - r_a, a the first station of s's block
//                              of 4 (fjl::station_dr)
//   coef              float    per frequency, fjl::coef_block floats:
//       itype 0: [nrp][nspec] A_s / sqrt(r_s) of each spectrum (J - Y: in
//                polar form, fjl::to_polar)
//       itype 1: [nrp][4 nspec] hat records per spectrum, then
//                [nrp][nspec] b_s and [nspec][nrp] kappa
//                (fjl::station_coef, fjl::coef_store)
//       then an 8-float header: the long-segment limit (itype 1) and the
//       anchor limit (fjl::coef_header)
//   T[nf][nc]         float    main-kernel results, one row per frequency
//   out[nc][nf]       float    (the reference's layout, src/FJcpu.cpp:247),
//                              formed from T by k_finish (optionally
//                              normalised per frequency)
//
// Launch: a workgroup is one wave of 64 lanes = 64 consecutive velocities
// of ONE frequency.  The station data of the loop are then wave-uniform and
// are read with scalar loads (one record per station per wave); each
// lane owns one (f,c) output and keeps the whole station loop in registers,
// so the main kernels are VALU-bound by construction.  Workgroups are
// numbered XCD-major so that the workgroups of one frequency share an XCD
// and its L2.  A wave's 64 results are consecutive velocities of one
// frequency, so they are stored into T[f][c] (one 256-byte segment per wave)
// rather than strided into out[c][f], where every lane would dirty its own
// cache line; k_finish transposes through LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fj_kernels.h"
#include "fj_lane.h"
#include "fj_main.h"
#include "fj_prep.h"

namespace fjk {

using namespace fjl;

// station arrays (fjl::station_entry), padded to ntot entries with the last
// distance and zero gaps
__global__ void k_stations(const float* __restrict__ r, int nr, int ntot, double* __restrict__ rr,
                           float* __restrict__ hh, float* __restrict__ ir,
                           float* __restrict__ isr, uint32_t* __restrict__ bw,
                           float* __restrict__ dr) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < ntot) {
    station_entry(r, nr, s, rr, hh, ir, isr, dr);
    bw[s] = station_block_word(r, nr, s);
  }
}

// Coefficients + [nr][nf] -> per-frequency records (fj_prep.h).
template <int ITYPE, int NSPEC, bool POLAR>
__global__ __launch_bounds__(256) void k_prep(const float* __restrict__ U0,
                                              const float* __restrict__ U1,
                                              const float* __restrict__ r,
                                              const float* __restrict__ f,
                                              const float* __restrict__ c, int nr, int nrp,
                                              int nf, int nc, float* __restrict__ coef) {
  prep_tile<ITYPE, NSPEC, POLAR>(U0, U1, r, f, c, nr, nrp, nf, nc, coef, blockIdx.x * 64,
                                 blockIdx.y * 64);
}

// Per-frequency scale for the normalisation (src/ccfj.pyx:225-226,
// out[:,i] /= max(abs(out[:,i]))), with Python's builtin-max semantics over
// the column: a NaN in row 0 propagates; a later NaN never compares greater
// and is skipped (fmaxf skips it too).  Block (f, b) reduces row f of
// T[b] (contiguous), 256 lanes, then wave shuffles.
__global__ __launch_bounds__(256) void k_rowmax(const float* __restrict__ T, int nf, int nc,
                                                float* __restrict__ m) {
  __shared__ float red[4];
  const int f = blockIdx.x, b = blockIdx.y;
  const float* row = T + ((size_t)b * nf + f) * nc;
  float v = 0.0f;   // |values| >= 0, and row 0 is always part of the max
  for (int j = threadIdx.x; j < nc; j += 256) v = fmaxf(v, fabsf(row[j]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float a0 = fabsf(row[0]);
    const float r = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    m[(size_t)b * nf + f] = a0 != a0 ? a0 : r;
  }
}

// out[b][c][f] = T[b][f][c] (/ m[b][f] when m is given): 64 x 64 tiles
// through LDS, loads coalesced along c, stores along f.  The division (not a
// reciprocal product) rounds as the reference's in-place division.
template <typename OutT>
__global__ __launch_bounds__(256) void k_finish(const float* __restrict__ T,
                                                const float* __restrict__ m, int nf, int nc,
                                                OutT* __restrict__ out) {
  __shared__ float tile[64][65];
  const int b = blockIdx.z;
  const int c0 = blockIdx.x * 64, f0 = blockIdx.y * 64;
  const size_t plane = (size_t)nf * nc;
  const float* Tb = T + b * plane;
  OutT* ob = out + b * plane;
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  for (int fl = grp; fl < 64; fl += 4) {
    const int f = f0 + fl, c = c0 + lane;
    if (f < nf && c < nc) tile[fl][lane] = Tb[(size_t)f * nc + c];
  }
  __syncthreads();
  const int f = f0 + lane;
  const float scale_src = (m && f < nf) ? m[(size_t)b * nf + f] : 1.0f;
  for (int cl = grp; cl < 64; cl += 4) {
    const int c = c0 + cl;
    if (f < nf && c < nc) {
      const float v = tile[lane][cl];
      ob[(size_t)c * nf + f] = (OutT)(m ? v / scale_src : v);
    }
  }
}

__global__ void k_fill(float* __restrict__ p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
Stations stations(const Problem& p) { return Stations{p.rr, p.hh, p.ir, p.isr, p.bw, p.dr}; }

template <int MODE>
static hipError_t launch_main(int itype, const Problem& p, hipStream_t st) {
  if (itype != 0) return launch_fj_linseg(MODE, p, st);   // fj_kernels_linseg.hip
  hipLaunchKernelGGL((k_fj<MODE, 0>), dim3((unsigned)main_blocks(p.nf, p.nc)), dim3(kBlock), 0,
                     st, stations(p), p.coef, p.f, p.c, p.nrp, p.nf, p.nc, p.T);
  return hipGetLastError();
}

template <int ITYPE, int NSPEC, bool POLAR = false>
static hipError_t launch_prep_t(const Problem& p, hipStream_t st) {
  dim3 grid((p.nf + 63) / 64, (p.nrp + 63) / 64);
  hipLaunchKernelGGL((k_prep<ITYPE, NSPEC, POLAR>), grid, dim3(256), 0, st, p.U0, p.U1, p.r, p.f, p.c,
                     p.nr, p.nrp, p.nf, p.nc, p.coef);
  return hipGetLastError();
}

size_t coef_floats(int mode, int itype, int nrp, int nf) {
  return (size_t)nf * coef_block(nrp, mode, itype);
}

int padded(int nr) { return padded_stations(nr); }

hipError_t launch_prep(int mode, int itype, const Problem& p, hipStream_t st) {
  const int ntot = p.nrp + kUnroll;
  hipLaunchKernelGGL(k_stations, dim3((ntot + 255) / 256), dim3(256), 0, st, p.r, p.nr, ntot,
                     p.rr, p.hh, p.ir, p.isr, p.bw, p.dr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const bool two = !(mode == MODE_J || mode == MODE_Y);
  if (polar_records(mode, itype)) return launch_prep_t<0, 2, true>(p, st);
  if (itype == 0) return two ? launch_prep_t<0, 2>(p, st) : launch_prep_t<0, 1>(p, st);
  return two ? launch_prep_t<1, 2>(p, st) : launch_prep_t<1, 1>(p, st);
}

hipError_t launch_fj(int mode, int itype, const Problem& p, hipStream_t st) {
  switch (mode) {
    case MODE_J: return launch_main<MODE_J>(itype, p, st);
    case MODE_Y: return launch_main<MODE_Y>(itype, p, st);
    case MODE_J_MINUS_Y: return launch_main<MODE_J_MINUS_Y>(itype, p, st);
    case MODE_ABS_J: return launch_main<MODE_ABS_J>(itype, p, st);
    case MODE_ABS_H: return launch_main<MODE_ABS_H>(itype, p, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_finish(const float* T, float* m, int nb, int nf, int nc, float* out,
                         hipStream_t st) {
  if (nb <= 0 || nf <= 0 || nc <= 0) return hipSuccess;
  if (m) {
    hipLaunchKernelGGL(k_rowmax, dim3(nf, nb), dim3(256), 0, st, T, nf, nc, m);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_finish<float>, dim3((nc + 63) / 64, (nf + 63) / 64, nb), dim3(256), 0, st,
                     T, (const float*)m, nf, nc, out);
  return hipGetLastError();
}

hipError_t launch_finish_f64(const float* T, float* m, int nb, int nf, int nc, double* out,
                             bool rowmax, hipStream_t st) {
  if (nb <= 0 || nf <= 0 || nc <= 0) return hipSuccess;
  if (rowmax) {
    hipLaunchKernelGGL(k_rowmax, dim3(nf, nb), dim3(256), 0, st, T, nf, nc, m);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_finish<double>, dim3((nc + 63) / 64, (nf + 63) / 64, nb), dim3(256), 0,
                     st, T, (const float*)m, nf, nc, out);
  return hipGetLastError();
}

hipError_t launch_fill(float* p, size_t n, float v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

}  // namespace fjk
