// fj_synth.h -- synthetic training fields, labels, item-list F-J.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fj_lane.h"

namespace fjk {

constexpr int kMaxModes = 4;

hipError_t launch_synth_fields(const float* r, const float* f, const float* modes,
                               const float* sigma, int B, int nr, int nf, uint64_t seed, float* U,
                               hipStream_t st);
hipError_t launch_synth_labels(const float* cgrid, const float* f, const float* modes, int B,
                               int H, int nf, float halfwidth, float amin, float* Y,
                               hipStream_t st);
hipError_t launch_pool_abs_max_ranges(const float* in, int B, int N, int H, int W,
                                      const int* ranges, float* out, hipStream_t st);
hipError_t launch_pool_abs_max(const float* in, int B, int H, int S, int W, float* out,
                               hipStream_t st);
hipError_t launch_label_argmax(const float* X, int B, int H, int W, float* Y, hipStream_t st);
hipError_t launch_fj_items(int mode, int itype, const fjl::Stations& S, const float* coef,
                           const float* f, const float* c, const int* items, int nitems, int nr,
                           int nrp, int nf, int nc, float* out, hipStream_t st);

}  // namespace fjk
