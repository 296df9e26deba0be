// fj_cc.h -- cross-correlation kernels (fj_cc.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace fjk {

hipError_t launch_cc_onebit(float* x, size_t n, hipStream_t st);
hipError_t launch_cc_gather(const float* data, int nsta, int npts, int L, int fold, int steplen,
                            int w0, int cw, float* seg, hipStream_t st);
hipError_t launch_cc_spec(const void* spec, int rows, int nbins, int zhalf, int nf, int fstride,
                          int whiten, void* S, hipStream_t st);
hipError_t launch_cc_use(const int* startend, int nsta, int nshifts, int steplen, int fftlen,
                         unsigned char* use, hipStream_t st);
hipError_t launch_cc_count(const unsigned char* use, const int* pairs, int npairs, int nsta,
                           int nshifts, float* count, hipStream_t st);
hipError_t launch_cc_pairs(const void* S, const unsigned char* use, const int* pairs, int npairs,
                           int nsta, int nf, int w0, int cw, void* acc, hipStream_t st);
// whiten: S2 receives the whitened spectra in split fp16 form (re_hi | im_hi
// << 16, re_lo | im_lo << 16), 8 bytes a value as unwhitened float2 ones
hipError_t launch_cc_spec_t(const void* spec, int rows, int nsta, int nbins, int zhalf, int nf,
                            int fstride, int whiten, const unsigned char* use, int w0, void* S2,
                            int* nonfinite, hipStream_t st);
hipError_t launch_cc_pair_map(const int* pairs, int npairs, int nsta, int ntile, int* pidx,
                              unsigned char* dup, unsigned char* tiles, hipStream_t st);
// whiten: the spectra are whitened (|x| <= 1), and the product takes split
// fp16 MFMAs (k_cc_gram_split); else FP32 ones (k_cc_gram).  first: acc
// is not initialised; the listed pairs are stored instead of added to
hipError_t launch_cc_gram(const void* S2, int nsta, int nf, int cw, int ntile, int whiten,
                          const unsigned char* tiles, const int* pidx, const int* nonfinite,
                          int first, void* acc, hipStream_t st);
// whiten: S2 holds the split form of launch_cc_spec_t.  first: acc is
// not initialised; stores every pair that launch_cc_gram (first) leaves
// (duplicates, all pairs of a chunk with a non-finite value, zeros for a
// pair naming a station outside [0, nsta))
hipError_t launch_cc_pairs_t(const void* S2, int whiten, const unsigned char* use,
                             const int* pairs, int npairs, int nsta, int nf, int w0, int cw,
                             const unsigned char* dup, const int* nonfinite, int first,
                             void* acc, hipStream_t st);
hipError_t launch_cc_fill_i32(int* x, size_t n, int v, hipStream_t st);
hipError_t launch_cc_interleave(const float* re, const float* im, size_t n, void* acc,
                                hipStream_t st);
hipError_t launch_cc_finish(const void* acc, const float* count, int npairs, int nf, float* re,
                            float* im, hipStream_t st);

}  // namespace fjk
