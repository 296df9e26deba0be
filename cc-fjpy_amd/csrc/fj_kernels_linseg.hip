// fj_kernels_linseg.hip -- the dense linear-segment (itype 1) F-J kernels,
// k_fj<MODE, 1> (fj_main.h), in a translation unit of their own.
//
// The build compiles this file with machine-level loop-invariant code motion
// off (-mllvm -disable-machine-licm).  With it on, the compiler hoists the
// double-precision constants of the small-argument polynomials (x < 8, in a
// rarely taken branch) out of the station loop into vector registers, which
// then stay live through the whole loop: the J - Y kernel held 168 VGPRs at
// 3 waves per SIMD and spilled 116 bytes per lane (written at entry, read
// back in serial chains in the branch), the J kernel 80 VGPRs and 104
// bytes.  Without it the constants are formed where they are used: J - Y 62
// VGPRs, 7 waves, no scratch; J 43 VGPRs, 8 waves, no scratch.  Measured on
// the MI355X at 19,900 pairs x 4096 f x 3000 c: J - Y 392.5 -> 322.7 ms, J
// 222.5 -> 222.1 ms.  The trapezoid kernels (fj_kernels.hip) keep it on:
// without it they were 2% slower (J 96.7 -> 98.5 ms, J - Y 138.4 -> 140.3).
#include <hip/hip_runtime.h>

#include "fj_kernels.h"
#include "fj_main.h"

namespace fjk {

template <int MODE>
static hipError_t launch_linseg(const Problem& p, hipStream_t st) {
  hipLaunchKernelGGL((k_fj<MODE, 1>), dim3((unsigned)main_blocks(p.nf, p.nc)), dim3(kBlock), 0,
                     st, stations(p), p.coef, p.f, p.c, p.nrp, p.nf, p.nc, p.T);
  return hipGetLastError();
}

hipError_t launch_fj_linseg(int mode, const Problem& p, hipStream_t st) {
  switch (mode) {
    case MODE_J: return launch_linseg<MODE_J>(p, st);
    case MODE_Y: return launch_linseg<MODE_Y>(p, st);
    case MODE_J_MINUS_Y: return launch_linseg<MODE_J_MINUS_Y>(p, st);
    case MODE_ABS_J: return launch_linseg<MODE_ABS_J>(p, st);
    case MODE_ABS_H: return launch_linseg<MODE_ABS_H>(p, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace fjk
