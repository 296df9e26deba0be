// fj_main.h -- the dense F-J main kernel k_fj<MODE, ITYPE>, shared by the
// two translation units that instantiate it: fj_kernels.hip (trapezoid,
// ITYPE 0) and fj_kernels_linseg.hip (linear segments, ITYPE 1), which the
// build compiles with different code-motion options.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fj_lane.h"

namespace fjk {

using namespace fjl;

// Lanes per workgroup of the dense kernels: one wave.  Measured on the
// MI355X (19,900 pairs x 4096 f x 3000 c): trapezoid J 96.9 ms against 102.8
// with 128 lanes and 105.2 with 256, J - Y 139.5 / 148.9 / 155.1 ms; linear
// segments J 222.8 / 229.0 / 231.2, J - Y 396.0 / 408.5 / 412.3 ms.  A
// workgroup of several waves holds its slots until its slowest wave ends.
constexpr int kBlock = 64;

// Hardware places workgroup w on XCD (w mod 8); tile T = (w mod 8) * per +
// w / 8 keeps consecutive tiles (same frequency) on one XCD.  Returns false
// for the padding workgroups of the last round.
__device__ inline bool tile_of(int ntiles, int* tile) {
  const int w = blockIdx.x;
  const int per = (ntiles + 7) >> 3;
  const int T = (w & 7) * per + (w >> 3);
  *tile = T;
  return T < ntiles;
}

// Workgroups of a dense launch over nf x nc: whole rounds of 8 (one per XCD).
inline long main_blocks(int nf, int nc) {
  const long ntiles = (long)nf * ((nc + kBlock - 1) / kBlock);
  return ((ntiles + 7) / 8) * 8;
}

template <int MODE, int ITYPE>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(FjOcc<MODE, ITYPE>::waves, 8))) void k_fj(Stations S, const float* __restrict__ coef,
                                              const float* __restrict__ fv,
                                              const float* __restrict__ cv, int nrp, int nf,
                                              int nc, float* __restrict__ T_out) {
  const int ncb = (nc + kBlock - 1) / kBlock;
  int T;
  if (!tile_of(nf * ncb, &T)) return;
  const int fi = T / ncb;
  const int j = (T - fi * ncb) * kBlock + threadIdx.x;
  // lanes past the last velocity compute a copy of it (the lane code needs
  // every lane of the wave active) and store nothing
  const double q = (double)fv[fi] / (double)cv[j < nc ? j : nc - 1];
  const float* __restrict__ cf = coef + (size_t)fi * coef_block(nrp, MODE, ITYPE);
  float v;
  if (ITYPE == 0)
    v = trapz_lane<MODE>(S, cf, nrp, q);
  else
    v = linseg_lane<MODE>(S, cf, nrp, q);
  if (j < nc) T_out[(size_t)fi * nc + j] = v;
}

}  // namespace fjk
