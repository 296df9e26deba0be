// fj_kernels.h -- interface between the host library and the F-J kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/ccfj.h"
#include "fj_lane.h"

namespace fjk {

// Device pointers of one transform.  U1 is the second spectrum for the
// two-spectrum modes (Hilbert-transformed spectrum, or imaginary part).
struct Problem {
  const float* U0;
  const float* U1;
  const float* r;
  const float* f;
  const float* c;
  int nr, nf, nc;
  int nrp;       // station count padded to the kernels' block (fjl::padded_stations)
  double* rr;    // [nrp + 8] workspace
  float* hh;     // [nrp + 8] workspace
  float* ir;     // [nrp + 8] workspace, 1 / r
  float* isr;    // [nrp + 8] workspace, 1 / sqrt(r)
  uint32_t* bw;  // [nrp + 8] workspace, block words (fjl::station_block_word)
  float* dr;     // [nrp + 8] workspace, offsets from the block anchors (fjl::station_dr)
  float* coef;   // coef_floats(...) workspace
  float* T;      // [nf][nc] main-kernel results (k_finish forms [nc][nf])
};

fjl::Stations stations(const Problem& p);
size_t coef_floats(int mode, int itype, int nrp, int nf);  // coef workspace
int padded(int nr);  // station count padded to the kernels' block
hipError_t launch_prep(int mode, int itype, const Problem& p, hipStream_t st);
hipError_t launch_fj(int mode, int itype, const Problem& p, hipStream_t st);
// the itype 1 part of launch_fj (fj_kernels_linseg.hip)
hipError_t launch_fj_linseg(int mode, const Problem& p, hipStream_t st);
// out[b][c][f] = T[b][f][c], divided by the column max |.| when m (workspace
// [nb][nf]) is given; nb stacked problems of nf x nc.
hipError_t launch_finish(const float* T, float* m, int nb, int nf, int nc, float* out,
                         hipStream_t st);
// the same normalised transpose into float64 (MWFJ's output type); m is
// required (nb x nf floats): the column maxima are formed into it when
// rowmax, else it holds them already
hipError_t launch_finish_f64(const float* T, float* m, int nb, int nf, int nc, double* out,
                             bool rowmax, hipStream_t st);
// p[0..n) = v (a kernel, so that a captured graph holds no memset node)
hipError_t launch_fill(float* p, size_t n, float v, hipStream_t st);

}  // namespace fjk
