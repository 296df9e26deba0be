// fj_mwfj.h -- interface of the multi-window F-J kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace fjk {

struct MwfjDev {
  int nr, npts, nf, nc, nwin;
  int fstride, nbins;  // spec row length: npts/2 + 1 rFFT bins, or npts/2 (zhalf)
  int zhalf;           // spec rows hold the half-length complex DFT (even npts)
  int func;
  int half;            // K rows per basis half (nr rounded up to 32)
  int nrhs;            // 2 nwin rounded up to 32
  const double* u;     // [nr][npts] traces, rows sorted by r
  const int* n1;       // [nwin][nr] window descriptions (sorted station order)
  const int* n2;
  const int* dl;
  const int* xidx;     // [nwin][nr] row of xrows for degenerate windows, else -1
  const double* xrows; // [nx][npts] explicit window rows
  int nx;
  const float* r;      // [nr] sorted, float32 (the reference's cast)
  const double* rr;    // [half] the same in double, padded with the last distance
  const float* ir;     // [half] 1 / r
  const float* f;
  const float* c;
  double* xw;          // tapered traces of one chunk of windows
  void* spec;          // rFFT of the chunk, complex double
  float* B;            // [nf][Kp][nrhs] coefficients (itype 0)
  float* T;            // [nwin][nf][nc] |O| before normalisation (itype 0)
  float* M;            // [nwin][nf] column maxima (itype 0)
  float* Mp;           // [mwfj_groups(nc)][nwin][nf] their partials per 32 velocities
  unsigned int* bmax;  // [nf] bits of max |B| per frequency (itype 0; zeroed before
                       // the coefficient kernels, which raise it by atomicMax)
  double* out;         // [nwin][nc][nf]
};

hipError_t launch_window(const MwfjDev& d, int w0, int cw, hipStream_t st);
// the used bins of windows w0 .. w0+cw-1 by the Goertzel recurrence over
// each window's support: spec [cw][nr][nf] complex double (k_mwfj_goertzel)
hipError_t launch_mwfj_goertzel(const MwfjDev& d, int w0, int cw, const int* rowbad,
                                const double2* tw, double2* spec, hipStream_t st);
// tw[m] = (cos, sin)(2 pi m / n), m < n (device, n entries)
hipError_t launch_mwfj_twiddles(int n, double2* tw, hipStream_t st);
// rowbad[s] = 1 when trace s holds a non-finite sample (device, [nr])
hipError_t launch_mwfj_rowbad(const MwfjDev& d, int* rowbad, hipStream_t st);
// sum of the windows' support lengths into *sum (device; zeroed by the caller)
hipError_t launch_mwfj_support(const MwfjDev& d, unsigned long long* sum, hipStream_t st);
hipError_t launch_mwfj_stations(const MwfjDev& d, double* rr, float* ir, hipStream_t st);
// zeros in the padding of B (rows nr..half-1, columns 2 nwin..nrhs-1) that
// launch_mwfj_coef leaves; nothing to do when there is none
hipError_t launch_mwfj_bpad(const MwfjDev& d, hipStream_t st);
hipError_t launch_mwfj_coef(const MwfjDev& d, int w0, int cw, hipStream_t st);
hipError_t launch_mwfj_mfma(const MwfjDev& d, hipStream_t st);
// velocity groups of 32 of the contraction kernel's grid (partial maxima)
int mwfj_groups(int nc);
hipError_t launch_mwfj_finish(const MwfjDev& d, hipStream_t st);
hipError_t launch_mwfj_stage(const MwfjDev& d, int wl, float* re, float* im, hipStream_t st);
hipError_t launch_to_double(const float* a, size_t n, double* out, hipStream_t st);

}  // namespace fjk
