// fj_batched.hip -- many independent sub-arrays of uneven size in one launch.
//
// No reference counterpart: the reference transforms one array per call
// (src/ccfj.pyx:202-227, one FJ/FH call each).  Here nb sub-arrays that
// share the frequency and velocity grids are prepared by two batched prep
// kernels and evaluated by ONE persistent kernel: a grid sized to the
// device's resident capacity pulls work items (sub-array, frequency,
// velocity tile) from an atomic counter, largest sub-arrays first, so the
// tail of the launch is made of the cheapest items.  The per-(f,c) work is
// the same lane code as the single-array kernels (fj_lane.h).
//
// The build compiles this file without machine-level loop-invariant code
// motion (as fj_kernels_linseg.hip): with it, the linear-segment J kernel
// held the small-argument polynomial constants in vector registers through
// the station loop and spilled 96 bytes per lane, stored again for every
// work item (4.3e10 bytes written and 1.5e11 read per launch of 4096 ragged
// sub-arrays x 512 f x 1024 c); without it, 43 VGPRs and no scratch, in
// the same time (604 against 602 ms).
//
// Layout (device):
//   U0/U1  f32 [sum nr_b][nf]        rows of all sub-arrays, each sorted by r
//   r      f32 [sum nr_b]
//   rr, hh, ir, isr, bw, dr  per sub-array b at poff[b], nrp_b + 8 entries each (padded)
//   coef   per sub-array b at coff[b]: nf frequencies of nrp_b records (fj_kernels.hip)
//   T      f32 [nb][nf][nc]     persistent-kernel results (coalesced along c)
//   out    f32 [nb][nc][nf]     formed from T by k_finish (fj_kernels.hip)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fj_batched.h"
#include "fj_lane.h"
#include "fj_prep.h"

namespace fjk {

using namespace fjl;

__global__ void k_stations_b(BatchDev d) {
  const int b = blockIdx.y;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t s0 = d.soff[b];
  const int nr = (int)(d.soff[b + 1] - s0);
  const int ntot = d.nrp[b] + kUnroll;
  if (s >= ntot || nr <= 0) return;
  const int64_t o = d.poff[b];
  station_entry(d.r + s0, nr, s, d.rr + o, d.hh + o, d.ir + o, d.isr + o, d.dr + o);
  d.bw[o + s] = station_block_word(d.r + s0, nr, s);
}

template <int ITYPE, int NSPEC, bool POLAR>
__global__ __launch_bounds__(256) void k_prep_b(BatchDev d) {
  const int b = blockIdx.z;
  const int64_t s00 = d.soff[b];
  const int nr = (int)(d.soff[b + 1] - s00);
  const int nrp = d.nrp[b];
  const int s0 = blockIdx.y * 64;
  if (s0 >= nrp || nr <= 0) return;   // uniform per block: no barrier skipped partially
  prep_tile<ITYPE, NSPEC, POLAR>(d.U0 + s00 * d.nf, d.U1 + s00 * d.nf, d.r + s00, d.f, d.c, nr, nrp,
                          d.nf, d.nc, d.coef + d.coff[b], blockIdx.x * 64, s0);
}

__global__ void k_reset_counter(unsigned int* counter, unsigned int start) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = start;
}

// Persistent F-J over work items (p, fi, cb): p indexes order (sub-arrays
// by descending size), cb a tile of TILE velocities.  Items [start, limit)
// are processed (the counter starts at start); every workgroup leaves when
// the counter passes limit.  The
// arrays come as separate __restrict__ parameters (not through BatchDev):
// with no store that may alias them, their uniform loads -- the station
// data of the loop -- are scalar loads, as in k_fj.
template <int MODE, int ITYPE, int TILE>
__global__ __launch_bounds__(TILE)
__attribute__((amdgpu_waves_per_eu(FjOcc<MODE, ITYPE>::waves, 8))) void k_fj_b(
    int nf, int nc, unsigned int limit, const int64_t* __restrict__ soff,
    const int64_t* __restrict__ poff, const int64_t* __restrict__ coff,
    const int* __restrict__ nrpa, const int* __restrict__ order, const float* __restrict__ fv,
    const float* __restrict__ cv, const double* __restrict__ rr, const float* __restrict__ hh,
    const float* __restrict__ ir, const float* __restrict__ isr, const uint32_t* __restrict__ bw,
    const float* __restrict__ dr, const float* __restrict__ coef, float* __restrict__ T,
    unsigned int* __restrict__ counter) {
  __shared__ unsigned int item_s;
  const unsigned int ncb = (unsigned int)((nc + TILE - 1) / TILE);
  while (true) {
    if (threadIdx.x == 0) item_s = atomicAdd(counter, 1u);
    __syncthreads();
    const unsigned int item = __builtin_amdgcn_readfirstlane(item_s);   // uniform
    __syncthreads();   // every lane has read item_s before it is overwritten
    if (item >= limit) break;
    const unsigned int cb = item % ncb;
    const unsigned int pf = item / ncb;
    const int fi = (int)(pf % (unsigned int)nf);
    const int b = order[pf / (unsigned int)nf];
    const int j = (int)cb * TILE + (int)threadIdx.x;
    const int64_t s0 = soff[b];
    const int nr = (int)(soff[b + 1] - s0);
    if (nr > 1) {   // uniform over the workgroup: every lane runs the lane code
      const int nrp = nrpa[b];
      const double q = (double)fv[fi] / (double)cv[j < nc ? j : nc - 1];
      const float* __restrict__ cf = coef + coff[b] + (size_t)fi * coef_block(nrp, MODE, ITYPE);
      const int64_t o = poff[b];
      const Stations S{rr + o, hh + o, ir + o, isr + o, bw + o, dr + o};
      float v;
      if (ITYPE == 0)
        v = trapz_lane<MODE>(S, cf, nrp, q);
      else
        v = linseg_lane<MODE, true>(S, cf, nrp, q);
      if (j < nc) T[((size_t)b * nf + fi) * nc + j] = v;
    } else if (j < nc) {   // fewer than two stations: empty sum, as the reference
      T[((size_t)b * nf + fi) * nc + j] = 0.0f;
    }
  }
}

// ---------------------------------------------------------------------------
template <int ITYPE, int NSPEC, bool POLAR = false>
static hipError_t prep_t(const BatchDev& d, int max_nrp, hipStream_t st) {
  dim3 grid((d.nf + 63) / 64, (max_nrp + 63) / 64, d.nb);
  hipLaunchKernelGGL((k_prep_b<ITYPE, NSPEC, POLAR>), grid, dim3(256), 0, st, d);
  return hipGetLastError();
}

hipError_t launch_prep_batched(int mode, int itype, const BatchDev& d, int max_nrp,
                               hipStream_t st) {
  const int mt = max_nrp + kUnroll;
  hipLaunchKernelGGL(k_stations_b, dim3((mt + 255) / 256, d.nb), dim3(256), 0, st, d);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const bool two = !(mode == MODE_J || mode == MODE_Y);
  if (polar_records(mode, itype)) return prep_t<0, 2, true>(d, max_nrp, st);
  if (itype == 0) return two ? prep_t<0, 2>(d, max_nrp, st) : prep_t<0, 1>(d, max_nrp, st);
  return two ? prep_t<1, 2>(d, max_nrp, st) : prep_t<1, 1>(d, max_nrp, st);
}

template <int MODE, int ITYPE, int TILE>
static const void* kernel_ptr() {
  return reinterpret_cast<const void*>(&k_fj_b<MODE, ITYPE, TILE>);
}

template <int MODE, int ITYPE>
static const void* pick(int tile) {
  switch (tile) {
    case 64: return kernel_ptr<MODE, ITYPE, 64>();
    case 128: return kernel_ptr<MODE, ITYPE, 128>();
    default: return kernel_ptr<MODE, ITYPE, 256>();
  }
}

const void* batched_kernel(int mode, int itype, int tile) {
  switch (mode * 2 + itype) {
    case MODE_J * 2 + 0: return pick<MODE_J, 0>(tile);
    case MODE_J * 2 + 1: return pick<MODE_J, 1>(tile);
    case MODE_Y * 2 + 0: return pick<MODE_Y, 0>(tile);
    case MODE_Y * 2 + 1: return pick<MODE_Y, 1>(tile);
    case MODE_J_MINUS_Y * 2 + 0: return pick<MODE_J_MINUS_Y, 0>(tile);
    case MODE_J_MINUS_Y * 2 + 1: return pick<MODE_J_MINUS_Y, 1>(tile);
    case MODE_ABS_J * 2 + 0: return pick<MODE_ABS_J, 0>(tile);
    case MODE_ABS_J * 2 + 1: return pick<MODE_ABS_J, 1>(tile);
    case MODE_ABS_H * 2 + 0: return pick<MODE_ABS_H, 0>(tile);
    case MODE_ABS_H * 2 + 1: return pick<MODE_ABS_H, 1>(tile);
  }
  return nullptr;
}

hipError_t launch_batched(const void* kern, int tile, int grid, const BatchDev& d,
                          unsigned int start, unsigned int limit, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, d.counter, start);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  int nf = d.nf, nc = d.nc;
  const int64_t *soff = d.soff, *poff = d.poff, *coff = d.coff;
  const int *nrp = d.nrp, *order = d.order;
  const float *f = d.f, *c = d.c, *hh = d.hh, *ir = d.ir, *isr = d.isr, *dr = d.dr,
              *coef = d.coef;
  const double* rr = d.rr;
  const uint32_t* bw = d.bw;
  float* T = d.T;
  unsigned int* counter = d.counter;
  void* args[] = {&nf, &nc, &limit, &soff, &poff, &coff, &nrp, &order, &f, &c,
                  &rr, &hh, &ir, &isr, &bw, &dr, &coef, &T, &counter};
  return hipLaunchKernel(kern, dim3(grid), dim3(tile), args, 0, st);
}

}  // namespace fjk
