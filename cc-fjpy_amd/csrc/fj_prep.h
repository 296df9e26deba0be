// fj_prep.h -- coefficient preparation on the device, shared by the
// single-array prep kernel (fj_kernels.hip) and the batched one
// (fj_batched.hip).
//
// One workgroup of 256 lanes turns a tile of 64 stations x 64 frequencies of
// the [nr][nf] spectra into the per-(frequency, station) records of
// fjl::station_coef, stored in the fjl::coef_store layout.  The tile and one
// halo station on each side (the records of itype 1 need the neighbours'
// values) pass through LDS: the loads are coalesced along f, the stores
// along s.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "fj_lane.h"

namespace fjk {

// U0/U1 [nr][nf] (U1 read when NSPEC = 2), r [nr] sorted, fgrid [nf], cgrid [nc]
// (the velocity grid the records will be used with: itype 1 leaves long
// segments out by min |c|); coef: nf blocks of fjl::coef_block floats; tile
// origin (f0, s0).  Stations nr..nrp-1 (padding) get zero records; the
// tiles with s0 = 0 write the header of their frequencies (the long-segment
// limit of itype 1, the anchor limit).
// POLAR: itype 0, J - Y records in polar form (fjl::to_polar).
template <int ITYPE, int NSPEC, bool POLAR = false>
__device__ inline void prep_tile(const float* __restrict__ U0, const float* __restrict__ U1,
                                 const float* __restrict__ r, const float* __restrict__ fgrid,
                                 const float* __restrict__ cgrid, int nr, int nrp, int nf, int nc,
                                 float* __restrict__ coef, int f0, int s0) {
  constexpr int TS = 64, TF = 64, LD = TF + 1, ROWS = TS + 2;
  constexpr int NREC = ITYPE == 0 ? NSPEC : 6 * NSPEC;
  const size_t block = fjl::coef_block(nrp, NSPEC == 1 ? fjl::MODE_J : fjl::MODE_J_MINUS_Y, ITYPE);
  __shared__ float tile[NSPEC][ROWS][LD];
  __shared__ float hlim_s[TF];
  __shared__ float red[4];
  const int lane = threadIdx.x & 63;
  const int grp = threadIdx.x >> 6;
  {
    // min |c| over the grid (exact, so every tile finds the same value)
    float m = INFINITY;
    for (int j = threadIdx.x; j < nc; j += 256) m = fminf(m, fabsf(cgrid[j]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fminf(m, __shfl_xor(m, o));
    if (lane == 0) red[grp] = m;
    __syncthreads();
    const float cmin = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
    if (threadIdx.x < TF) {
      const int fi = f0 + threadIdx.x;
      const float h =
          ITYPE == 1 && fi < nf ? fjl::long_segment_limit(fgrid[fi], cmin) : 0.0f;
      hlim_s[threadIdx.x] = h;
      if (s0 == 0 && fi < nf) {
        float* hd = coef + (size_t)fi * block + (size_t)nrp * NREC;
        hd[0] = h;
        hd[1] = fjl::anchor_limit(fgrid[fi], cmin);
      }
    }
  }
  for (int row = grp; row < ROWS; row += 4) {   // stations s0 - 1 .. s0 + TS
    const int s = s0 - 1 + row;
    const int f = f0 + lane;
    const bool ok = s >= 0 && s < nr && f < nf;
    const size_t idx = ok ? (size_t)s * nf + f : 0;
    tile[0][row][lane] = ok ? U0[idx] : 0.0f;
    if (NSPEC > 1) tile[NSPEC - 1][row][lane] = ok ? U1[idx] : 0.0f;
  }
  __syncthreads();
  for (int it = threadIdx.x; it < TS * TF; it += 256) {
    const int sl = it & (TS - 1);
    const int fl = it >> 6;
    const int s = s0 + sl, f = f0 + fl;
    if (s >= nrp || f >= nf) continue;
    float rec[NREC];
    if (s >= nr) {
#pragma unroll
      for (int k = 0; k < NREC; k++) rec[k] = 0.0f;
    } else {
      float gp[NSPEC], gc[NSPEC], gn[NSPEC];
#pragma unroll
      for (int k = 0; k < NSPEC; k++) {
        gp[k] = tile[k][sl][fl];
        gc[k] = tile[k][sl + 1][fl];
        gn[k] = tile[k][sl + 2][fl];
      }
      const float rp = s > 0 ? r[s - 1] : 0.0f;
      const float rn = s + 1 < nr ? r[s + 1] : 0.0f;
      fjl::station_coef<ITYPE, NSPEC>(s, nr, rp, r[s], rn, gp, gc, gn,
                                      ITYPE == 1 ? hlim_s[fl] : 0.0f, rec);
      if (POLAR) fjl::to_polar(rec);
    }
    fjl::coef_store<ITYPE, NSPEC>(coef + (size_t)f * block, nrp, s, rec);
  }
}

}  // namespace fjk
