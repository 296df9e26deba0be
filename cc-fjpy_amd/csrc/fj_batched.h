// fj_batched.h -- interface of the batched (many sub-array) F-J kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ccfj.h"

namespace fjk {

// Device view of a batch; every pointer is a device pointer.  Passed to the
// kernels by value.
struct BatchDev {
  int nb, nf, nc;
  const int64_t* soff;   // [nb+1] first station (row of U, entry of r) per sub-array
  const int64_t* poff;   // [nb]   offset of the padded station arrays (rr, hh)
  const int64_t* coff;   // [nb]   offset (floats) of the coefficient block
  const int* nrp;        // [nb]   padded station count
  const int* order;      // [nb]   sub-arrays by descending nrp (work order)
  const float* U0;
  const float* U1;
  const float* r;
  const float* f;
  const float* c;
  double* rr;
  float* hh;
  float* ir;
  float* isr;
  uint32_t* bw;          // block words (fjl::station_block_word), at poff like rr
  float* dr;             // offsets from the block anchors (fjl::station_dr), at poff
  float* coef;
  float* T;              // [nb][nf][nc] persistent-kernel results
  float* m;              // [nb][nf] column scales (normalisation)
  float* out;            // [nb][nc][nf]
  unsigned int* counter; // work-queue head
};

hipError_t launch_prep_batched(int mode, int itype, const BatchDev& d, int max_nrp,
                               hipStream_t st);
const void* batched_kernel(int mode, int itype, int tile);
// items [start, limit) of the work queue
hipError_t launch_batched(const void* kern, int tile, int grid, const BatchDev& d,
                          unsigned int start, unsigned int limit, hipStream_t st);

}  // namespace fjk
