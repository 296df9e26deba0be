// fj_cc.hip -- ambient-noise cross-correlation on the device: the producer
// of the spectra fj_noise consumes.
//
// What is computed (reference: src/CrossCorr.cpp:33-159): the record of
// every station is cut into windows of fftlen samples every steplen
// samples; each window's real FFT is optionally whitened (divided by
// max(|X|, 1e-8)); for every station pair (a, b) the products
// conj(X_a) X_b at bins j*fstride (j < nf) are summed over the windows in
// which both stations have data, then divided by that count.
//
// Layout (device):
//   data  f32 [nsta][npts]            (one-bit: replaced by its sign first)
//   seg   f32 [cw][nsta][L]           windows of a chunk, gathered (L = fftlen,
//                                     or fftlen / fstride when folded)
//   spec  c64 [cw][nsta][L/2 (+1)]    their rFFT (hipFFT, single precision): for
//                                     even L the length-L/2 complex DFT of
//                                     the packed samples, bins formed on read
//   S     c64 [cw][nsta][nf]          bins j*fstride, whitened
//   acc   c64 [npairs][nf]            running sums over windows
//   use   u8  [nshifts][nsta]         station has the whole window
//   count f32 [npairs]                windows both stations of a pair hold
//
// The pair kernel reads S rows of two stations per pair; a chunk's S
// (cw * nsta * nf * 8 bytes) stays in L2/MALL while all pairs sweep it, and
// acc is read and written once per chunk of cw windows, not per window.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include <type_traits>

#include "fj_cc.h"

namespace fjk {

__global__ void k_cc_onebit(float* __restrict__ x, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = x[i];
  x[i] = v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f);
}

// seg[wl][i][t] = sum over q < fold of data[i][(w0 + wl) * steplen + t + q L],
// t < L = fftlen / fold; one workgroup per window row (a workgroup per 256
// samples made 878k short workgroups for 231 windows of 200 stations,
// 0.70 ms, bound by their dispatch).  Rows are numbered station-major, (i,
// wl) -> i cw + wl, and split into 8 contiguous ranges, one per XCD
// (workgroup b runs on XCD b mod 8 and takes row (b mod 8) per + b / 8):
// the overlapping windows of a station run on one XCD and share its L2.
// Numbered window-fastest across the XCDs, each XCD read nearly every
// record (90% overlap, 8 XCDs).  V = 4: 16-byte stores and loads, a window
// start off the 16-byte grid read as two aligned loads and a shift (SH =
// start mod 4, uniform over the workgroup); V = 1: one sample a step.
// fold = 1 is the plain window.  Folding by the bin stride s (when s
// divides fftlen) is exact for the bins that are kept: X[s m] = sum_n
// (sum_q x[n + q L]) e^{-2 pi i m n / L}, so an rFFT of length L gives bins
// 0, s, 2s, ... of the length-fftlen one.
template <int SH>
__device__ inline float4 load_shifted(const float4* __restrict__ xb, int t) {
  const float4 A = xb[t];
  if (SH == 0) return A;
  // the last sample wanted, xb[t + 1] element SH - 1, is in the window, so
  // the whole aligned 16 bytes of xb[t + 1] are in the array
  const float4 B = xb[t + 1];
  if (SH == 1) return make_float4(A.y, A.z, A.w, B.x);
  if (SH == 2) return make_float4(A.z, A.w, B.x, B.y);
  return make_float4(A.w, B.x, B.y, B.z);
}

template <int SH>
__device__ inline void gather_row4(const float4* __restrict__ xb, float4* __restrict__ y, int n,
                                   int fold) {
#pragma unroll 4
  for (int t = threadIdx.x; t < n; t += 256) {
    float4 v = load_shifted<SH>(xb, t);
    for (int q = 1; q < fold; q++) {
      const float4 u = load_shifted<SH>(xb, t + q * n);
      v.x += u.x;
      v.y += u.y;
      v.z += u.z;
      v.w += u.w;
    }
    y[t] = v;
  }
}

template <int V>
__global__ __launch_bounds__(256) void k_cc_gather(const float* __restrict__ data, int npts,
                                                   int L, int fold, int steplen, int w0, int cw,
                                                   int nsta, float* __restrict__ seg) {
  const int per = gridDim.x >> 3;
  const long long row = (long long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (row >= (long long)cw * nsta) return;         // uniform over the block
  const int i = (int)(row / cw), wl = (int)(row - (long long)i * cw);
  const size_t s0 = (size_t)i * npts + (size_t)(w0 + wl) * steplen;
  float* y = seg + ((size_t)wl * nsta + i) * L;
  if (V == 4) {
    const int sh = (int)(s0 & 3);
    const float4* xb = reinterpret_cast<const float4*>(data + (s0 - sh));
    float4* y4 = reinterpret_cast<float4*>(y);
    const int n = L / 4;
    switch (sh) {   // uniform
      case 0: gather_row4<0>(xb, y4, n, fold); break;
      case 1: gather_row4<1>(xb, y4, n, fold); break;
      case 2: gather_row4<2>(xb, y4, n, fold); break;
      default: gather_row4<3>(xb, y4, n, fold); break;
    }
  } else {
    const float* x = data + s0;
#pragma unroll 4
    for (int t = threadIdx.x; t < L; t += 256) {
      float v = x[t];
      for (int q = 1; q < fold; q++) v += x[t + q * L];
      y[t] = v;
    }
  }
}

// Bin k of the real DFT of a row of L = 2M samples.  zhalf = 0: row holds
// the rFFT itself (M + 1 bins).  zhalf = 1: row holds the length-M complex
// DFT Z of z[n] = x[2n] + i x[2n+1]; for k <= M
//   X[k] = E[k] + e^{-i pi k / M} O[k],
//   E[k] = (Z[k] + conj Z[M-k]) / 2,  O[k] = (Z[k] - conj Z[M-k]) / (2i),
// with Z[M] = Z[0]: only the bins that are used are formed.
__device__ inline float2 rfft_bin_f(const float2* __restrict__ row, int k, int M, int zhalf) {
  if (!zhalf) return row[k];
  const float2 a = row[k < M ? k : 0];
  const float2 b = row[k == 0 ? 0 : M - k];
  const float er = 0.5f * (a.x + b.x), ei = 0.5f * (a.y - b.y);
  const float orr = 0.5f * (a.y + b.y), oi = -0.5f * (a.x - b.x);
  float sn, cs;
  sincospif((float)k / (float)M, &sn, &cs);   // e^{-i pi k/M} = cs - i sn
  return make_float2(er + cs * orr + sn * oi, ei + cs * oi - sn * orr);
}

// S[row][j] = spec[row][j * fstride] (whitened); rows = cw * nsta
__global__ void k_cc_spec(const float2* __restrict__ spec, int nbins, int zhalf, int nf,
                          int fstride, int whiten, float2* __restrict__ S) {
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  const size_t row = blockIdx.x;
  if (j >= nf) return;
  float2 v = rfft_bin_f(spec + row * nbins, j * fstride, nbins, zhalf);
  if (whiten) {
    // as the reference: |X| in float, floored at 1e-8, then a division
    const float a = fmaxf(sqrtf(v.x * v.x + v.y * v.y), 1e-8f);
    v.x = v.x / a;
    v.y = v.y / a;
  }
  S[row * nf + j] = v;
}

// use[k][i] = station i holds window k: k*steplen >= start_i and
// k*steplen + fftlen < end_i (src/CrossCorr.cpp:91-98)
__global__ void k_cc_use(const int* __restrict__ startend, int nsta, size_t n, int steplen,
                         int fftlen, unsigned char* __restrict__ use) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int i = (int)(e % nsta);
  const long long t0 = (long long)(e / nsta) * steplen;
  use[e] = t0 >= startend[2 * i] && t0 + fftlen < startend[2 * i + 1];
}

// count[p] = windows that both stations of pair p hold (0 for a pair that
// names a station outside [0, nsta)).  One wave per pair, 64 windows a
// step counted by ballot (a thread per pair looping over the windows took
// 58 us for 19,900 pairs and 231 windows: 78 blocks).
__global__ __launch_bounds__(256) void k_cc_count(const unsigned char* __restrict__ use,
                                                  const int* __restrict__ pairs, int npairs,
                                                  int nsta, int nshifts, float* __restrict__ count) {
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (p >= npairs) return;   // uniform over the wave
  const int a = pairs[2 * p], b = pairs[2 * p + 1];
  int n = 0;
  if (a >= 0 && a < nsta && b >= 0 && b < nsta) {
    for (int k0 = 0; k0 < nshifts; k0 += 64) {   // uniform
      const int k = k0 + lane;
      const bool both = k < nshifts && (use[(size_t)k * nsta + a] & use[(size_t)k * nsta + b]);
      n += __builtin_popcountll(__builtin_amdgcn_ballot_w64(both));
    }
  }
  if (lane == 0) count[p] = (float)n;   // the reference divides the float sums by it
}

// acc[p][j] += sum over wl < cw with use[w0+wl][a] && use[w0+wl][b] of
// conj(S[wl][a][j]) S[wl][b][j].  Block: one pair (grid x), 256 bins (y).
// A pair naming a station outside [0, nsta) adds nothing.
__global__ __launch_bounds__(256) void k_cc_pairs(const float2* __restrict__ S,
                                                  const unsigned char* __restrict__ use,
                                                  const int* __restrict__ pairs, int nsta,
                                                  int nf, int w0, int cw,
                                                  float2* __restrict__ acc) {
  const int p = blockIdx.x;
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  const int a = pairs[2 * p], b = pairs[2 * p + 1];
  if (j >= nf || a < 0 || a >= nsta || b < 0 || b >= nsta) return;
  float re = 0.0f, im = 0.0f;
  for (int wl = 0; wl < cw; wl++) {
    const unsigned char* u = use + (size_t)(w0 + wl) * nsta;
    if (!(u[a] && u[b])) continue;   // uniform over the block
    const float2 xa = S[((size_t)wl * nsta + a) * nf + j];
    const float2 xb = S[((size_t)wl * nsta + b) * nf + j];
    re = __builtin_fmaf(xa.x, xb.x, __builtin_fmaf(xa.y, xb.y, re));
    im = __builtin_fmaf(xa.x, xb.y, __builtin_fmaf(-xa.y, xb.x, im));
  }
  float2* o = acc + (size_t)p * nf + j;
  float2 v = *o;
  v.x += re;
  v.y += im;
  *o = v;
}

// ---- Gram form of the pair stage -------------------------------------------
// For a dense pair list the pair sums are a Gram product per bin: with the
// spectra of the stations lacking window w set to zero, for every bin j
//   G_j[a][b] = sum_w conj(S[w][a][j]) S[w][b][j],
// which v_mfma_f32_32x32x2_f32 evaluates for a 32 x 32 station tile per
// wave: the real part with K = (re, im) of one window, the imaginary part
// with (re, -im) against (im, re).  Each (station pair, bin) is read once
// per tile instead of once per pair.  Exactness against the per-pair form:
// a pair listed twice, and a chunk holding a non-finite used spectrum value
// (whose product with a masked zero would give NaN where the reference
// skips the window), fall back to the per-pair kernel.

typedef _Float16 cch2 __attribute__((ext_vector_type(2)));

// A whitened value (|re|, |im| <= 1) in split fp16 form, 8 bytes as its
// float2: x = hi + lo, hi = x rounded to fp16, lo = x - hi rounded to fp16;
// .x = re_hi | im_hi << 16, .y = re_lo | im_lo << 16.  The pair products
// take these words as MFMA operands directly.
__device__ inline uint2 split16(float2 v) {
  const _Float16 rh = (_Float16)v.x, ih = (_Float16)v.y;
  const _Float16 rl = (_Float16)(v.x - (float)rh), il = (_Float16)(v.y - (float)ih);
  return make_uint2(__builtin_bit_cast(unsigned, cch2{rh, ih}),
                    __builtin_bit_cast(unsigned, cch2{rl, il}));
}

__device__ inline float2 join16(uint2 u) {
  const cch2 h = __builtin_bit_cast(cch2, u.x), l = __builtin_bit_cast(cch2, u.y);
  return make_float2((float)h[0] + (float)l[0], (float)h[1] + (float)l[1]);
}

// S2[j][wl][s] = spectrum bin j*fstride of window w0+wl, station s,
// whitened, zero where the station lacks the window; flags a non-finite
// used value.  32 x 32 (rows x bins) tiles through LDS.
// Whitened spectra are stored in the split form of split16.
__global__ __launch_bounds__(256) void k_cc_spec_t(const float2* __restrict__ spec, int rows,
                                                   int nsta, int nbins, int zhalf, int nf,
                                                   int fstride,
                                                   int whiten, const unsigned char* __restrict__ use,
                                                   int w0, float2* __restrict__ S2,
                                                   int* __restrict__ nonfinite) {
  __shared__ float2 t[32][33];
  const int r0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  bool bad = false;
  for (int k = ty; k < 32; k += 8) {
    const int row = r0 + k, j = j0 + tx;
    float2 v = make_float2(0.0f, 0.0f);
    if (row < rows && j < nf) {
      const int wl = row / nsta, st = row - wl * nsta;
      if (use[(size_t)(w0 + wl) * nsta + st]) {
        v = rfft_bin_f(spec + (size_t)row * nbins, j * fstride, nbins, zhalf);
        if (whiten) {
          const float a = fmaxf(sqrtf(v.x * v.x + v.y * v.y), 1e-8f);
          v.x = v.x / a;
          v.y = v.y / a;
        }
        bad |= !(isfinite(v.x) && isfinite(v.y));
      }
    }
    t[k][tx] = v;
  }
  if (__builtin_amdgcn_ballot_w64(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(nonfinite, 1);
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int j = j0 + k, row = r0 + tx;
    if (j < nf && row < rows) {
      if (whiten)
        reinterpret_cast<uint2*>(S2)[(size_t)j * rows + row] = split16(t[tx][k]);
      else
        S2[(size_t)j * rows + row] = t[tx][k];
    }
  }
}

// pidx[a][b] = p for the pairs of the list (first come; a later duplicate
// is marked in dup and left to the per-pair kernel); tiles[(a/32)][(b/32)]
// marks the 32 x 32 tiles holding a listed pair.  pidx starts at -1.
__global__ void k_cc_pair_map(const int* __restrict__ pairs, int npairs, int nsta, int ntile,
                              int* __restrict__ pidx, unsigned char* __restrict__ dup,
                              unsigned char* __restrict__ tiles) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int a = pairs[2 * p], b = pairs[2 * p + 1];
  unsigned char d = 0;
  if (a >= 0 && a < nsta && b >= 0 && b < nsta) {
    const int old = atomicCAS(&pidx[(size_t)a * nsta + b], -1, p);
    if (old != -1) {
      d = 1;
    } else {
      tiles[(a >> 5) * ntile + (b >> 5)] = 1;
    }
  }
  dup[p] = d;
}

typedef float ccx16 __attribute__((ext_vector_type(16)));

// windows per batch of the Gram loop (k_cc_gram).  Measured on the MI355X at
// 200 stations, 231 windows and 1000 bins: 0.93 ms at 8 (124 VGPRs), 0.95 at
// 4, 1.27 at 16 (188 VGPRs, 2 waves); 1.00 ms with four windows' loads
// waited for at each step.
constexpr int kGramU = 8;

// One wave per (32 x 32 station tile, bin); 4 waves (4 bins) per block.
// acc[p][j] += G_j[a][b] for the listed pairs of the tile.
// Workgroup numbering is XCD-major: workgroups go to the 8 XCDs in turn
// (w mod 8), and the m-th bin group slot of XCD x (workgroups k =
// m ntile^2 .. (m + 1) ntile^2 - 1 of that XCD, one per tile) takes bin
// group 4 (8 (m / 4) + x) + m % 4, so that every tile of a bin group runs
// on one XCD and reads the group's spectra through that XCD's L2, and the
// 16 bins of 4 consecutive groups (one 128-byte line of a pair's sums) are
// written through one L2.  Numbered tile-major within each bin group (all
// tiles of a group on consecutive workgroups, spread over the XCDs), each
// XCD fetched the spectra again: 3.0 GB per call at the 200-station size,
// 8x the spectra.  With one group per XCD in turn (4 bins, 32 bytes of a
// line), the split form took 0.64 ms against 0.52 ms for runs of 4, at
// 200 stations, 231 windows and 1000 bins, its sums stored on the first
// chunk.
// v_mfma_f32_32x32x2_f32 operand layout (lane l): A[i = l%32][k = l/32],
// B[k = l/32][n = l%32]; result register v holds row 8(v/4) + 4(l/32) + v%4,
// column l%32.
__global__ __launch_bounds__(256) void k_cc_gram(const float2* __restrict__ S2, int nsta, int nf,
                                                 int cw, int ntile,
                                                 const unsigned char* __restrict__ tiles,
                                                 const int* __restrict__ pidx,
                                                 const int* __restrict__ nonfinite,
                                                 int first, float2* __restrict__ acc) {
  const int nt2 = ntile * ntile;
  const int k = blockIdx.x >> 3, m = k / nt2, t = k % nt2;
  const int g = (((m >> 2) << 3) + (blockIdx.x & 7)) * 4 + (m & 3);
  if (g * 4 >= nf) return;                        // uniform over the block
  const int ta = t / ntile, tb = t - ta * ntile;
  if (!tiles[t] || *nonfinite) return;            // uniform over the block
  const int j = g * 4 + (threadIdx.x >> 6);
  if (j >= nf) return;                            // uniform over the wave
  const int lane = threadIdx.x & 63, i = lane & 31, kq = lane >> 5;
  const int a = ta * 32 + i, b = tb * 32 + i;
  const bool av = a < nsta, bv = b < nsta;
  const size_t rows = (size_t)cw * nsta;
  const float2* Sj = S2 + (size_t)j * rows;
  ccx16 gre, gim;
#pragma unroll
  for (int v = 0; v < 16; v++) {
    gre[v] = 0.0f;
    gim[v] = 0.0f;
  }
  // windows in batches of kGramU, the next batch's loads issued before the
  // current batch's MFMAs so that their latency overlaps the matrix work
  // (windows past cw load zeros, which add nothing)
  auto load = [&](int w, float2& xa, float2& xb) {
    const bool ok = w < cw;
    const float2* Sw = Sj + (size_t)(ok ? w : 0) * nsta;
    xa = ok && av ? Sw[a] : make_float2(0.0f, 0.0f);
    xb = ok && bv ? Sw[b] : make_float2(0.0f, 0.0f);
  };
  float2 xa[kGramU], xb[kGramU];
#pragma unroll
  for (int u = 0; u < kGramU; u++) load(u, xa[u], xb[u]);
  for (int w0 = 0; w0 < cw; w0 += kGramU) {
    float2 na[kGramU], nb[kGramU];
    const bool more = w0 + kGramU < cw;   // uniform
    if (more) {
#pragma unroll
      for (int u = 0; u < kGramU; u++) load(w0 + kGramU + u, na[u], nb[u]);
    }
#pragma unroll
    for (int u = 0; u < kGramU; u++) {
      // real: ar br + ai bi;  imaginary: ar bi + (-ai) br
      const float a1 = kq ? xa[u].y : xa[u].x, b1 = kq ? xb[u].y : xb[u].x;
      const float a2 = kq ? -xa[u].y : xa[u].x, b2 = kq ? xb[u].x : xb[u].y;
      gre = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, gre, 0, 0, 0);
      gim = __builtin_amdgcn_mfma_f32_32x32x2f32(a2, b2, gim, 0, 0, 0);
    }
    if (more) {
#pragma unroll
      for (int u = 0; u < kGramU; u++) {
        xa[u] = na[u];
        xb[u] = nb[u];
      }
    }
  }
  if (!bv) return;
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int ar = ta * 32 + 8 * (v >> 2) + 4 * kq + (v & 3);
    if (ar >= nsta) continue;
    const int p = pidx[(size_t)ar * nsta + b];
    if (p < 0) continue;
    float2* o = acc + (size_t)p * nf + j;
    float2 s = first ? make_float2(0.0f, 0.0f) : *o;   // first: nothing to add to
    s.x += gre[v];
    s.y += gim[v];
    *o = s;
  }
}

// The Gram product of whitened spectra (|x| <= 1) in split fp16 products
// on v_mfma_f32_32x32x16_f16: each value held as hi = x rounded to fp16 and
// lo = the rounded remainder (22 significand bits together; split16, written
// by k_cc_spec_t), each product as hi hi + hi lo + lo hi with FP32
// accumulation (about 2^-22 of |a||b| per term).  K = 16 holds 8 windows x
// (re, im): lane l takes station rows i = l % 32 and windows 4 (l / 32) +
// 0..3 of each batch of 8, A[i][k] = (ar, ai) of its windows; the real part
// takes B[k][n] = (br, bi), the imaginary part (bi, -br) against the same A.
// The stored words are the operands: A and the real part's B as loaded, the
// imaginary part's B a half swap and a sign flip per word.  Six 32x32x16
// MFMAs per 8 windows against sixteen 32x32x2 FP32 ones, and each lane loads
// only its own windows; two batches in flight, the next one's loads issued
// before the current one's products.  Same numbering, tiles and result
// layout as k_cc_gram.
typedef _Float16 cch8 __attribute__((ext_vector_type(8)));
typedef unsigned ccu4 __attribute__((ext_vector_type(4)));

// At least 6 waves per SIMD (at most 80 registers, no spills): 0.46-0.48 ms
// against 0.52 at the 5 its 94 registers allow unbounded, 0.59 at 8 (spills
// to scratch), and 0.49 at 6 with each batch's loads waited for at once.
__global__ __launch_bounds__(256, 6) void k_cc_gram_split(const uint2* __restrict__ S2, int nsta,
                                                       int nf, int cw, int ntile,
                                                       const unsigned char* __restrict__ tiles,
                                                       const int* __restrict__ pidx,
                                                       const int* __restrict__ nonfinite,
                                                       int first, float2* __restrict__ acc) {
  const int nt2 = ntile * ntile;
  const int k = blockIdx.x >> 3, m = k / nt2, t = k % nt2;   // numbered as k_cc_gram
  const int g = (((m >> 2) << 3) + (blockIdx.x & 7)) * 4 + (m & 3);
  if (g * 4 >= nf) return;                        // uniform over the block
  const int ta = t / ntile, tb = t - ta * ntile;
  if (!tiles[t] || *nonfinite) return;            // uniform over the block
  const int j = g * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (j >= nf) return;                            // uniform over the wave
  const int lane = threadIdx.x & 63, i = lane & 31, kq = lane >> 5;
  const int a = ta * 32 + i, b = tb * 32 + i;
  const bool bv = b < nsta;
  // rows and columns of stations past nsta are never stored: they read
  // the last station's values instead of branching round their loads
  const int ac = min(a, nsta - 1), bc = min(b, nsta - 1);
  const uint2* Sj = S2 + (size_t)j * cw * nsta;   // uniform over the wave
  ccx16 gre, gim;
#pragma unroll
  for (int v = 0; v < 16; v++) {
    gre[v] = 0.0f;
    gim[v] = 0.0f;
  }
  // this lane's 4 windows of the batch from w0; a window past cw reads
  // window cw - 1 and its A operand is zeroed, so that it adds nothing
  const int last = (cw - 1) * nsta;
  auto load = [&](int w0, uint2* xa, uint2* xb) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int wq = 4 * kq + u;                  // the window's place in the batch
      const bool ok = wq < cw - w0;
      const int o = min(w0 * nsta + wq * nsta, last);
      const uint2 va = Sj[(unsigned)(o + ac)];
      xa[u] = make_uint2(ok ? va.x : 0u, ok ? va.y : 0u);
      xb[u] = Sj[(unsigned)(o + bc)];
    }
  };
  // (re, im) word -> (im, -re)
  auto conj_swap = [](unsigned x) { return __builtin_amdgcn_alignbit(x, x, 16) ^ 0x80000000u; };
  auto batch = [&](const uint2* xa, const uint2* xb) {
    const cch8 ah = __builtin_bit_cast(cch8, ccu4{xa[0].x, xa[1].x, xa[2].x, xa[3].x});
    const cch8 al = __builtin_bit_cast(cch8, ccu4{xa[0].y, xa[1].y, xa[2].y, xa[3].y});
    const cch8 bh = __builtin_bit_cast(cch8, ccu4{xb[0].x, xb[1].x, xb[2].x, xb[3].x});
    const cch8 bl = __builtin_bit_cast(cch8, ccu4{xb[0].y, xb[1].y, xb[2].y, xb[3].y});
    const cch8 ch = __builtin_bit_cast(cch8, ccu4{conj_swap(xb[0].x), conj_swap(xb[1].x),
                                                  conj_swap(xb[2].x), conj_swap(xb[3].x)});
    const cch8 cl = __builtin_bit_cast(cch8, ccu4{conj_swap(xb[0].y), conj_swap(xb[1].y),
                                                  conj_swap(xb[2].y), conj_swap(xb[3].y)});
    gre = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, gre, 0, 0, 0);
    gim = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, ch, gim, 0, 0, 0);
    gre = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, gre, 0, 0, 0);
    gim = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, cl, gim, 0, 0, 0);
    gre = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, gre, 0, 0, 0);
    gim = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, ch, gim, 0, 0, 0);
  };
  // batches in pairs; a batch wholly past cw adds nothing
  uint2 xa[4], xb[4], na[4], nb[4];
  load(0, xa, xb);
  for (int w0 = 0; w0 < cw; w0 += 16) {           // uniform
    load(w0 + 8, na, nb);
    batch(xa, xb);
    load(w0 + 16, xa, xb);
    batch(na, nb);
  }
  if (!bv) return;
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int ar = ta * 32 + 8 * (v >> 2) + 4 * kq + (v & 3);
    if (ar >= nsta) continue;
    const int p = pidx[(size_t)ar * nsta + b];
    if (p < 0) continue;
    float2* o = acc + (size_t)p * nf + j;
    float2 s = first ? make_float2(0.0f, 0.0f) : *o;   // first: nothing to add to
    s.x += gre[v];
    s.y += gim[v];
    *o = s;
  }
}

// The per-pair form on the S2 layout, for what the Gram form leaves: every
// pair when the chunk holds a non-finite used value, else the duplicates
// (and, on the first chunk, zeros for pairs naming a station outside
// [0, nsta), so that the sums need no fill).
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_cc_pairs_t(const float2* __restrict__ S2,
                                                    const unsigned char* __restrict__ use,
                                                    const int* __restrict__ pairs, int npairs,
                                                    int nsta, int nf, int w0, int cw,
                                                    const unsigned char* __restrict__ dup,
                                                    const int* __restrict__ nonfinite, int first,
                                                    float2* __restrict__ acc) {
  // block: 64 pairs (x) x 256 bins (y); the pairs classified first, one
  // thread each: 0 nothing to do, 1 summed here, 2 stored as zeros (first,
  // a station outside [0, nsta)).  First: every pair the Gram form does not
  // store is stored here.
  __shared__ int sa[64], sb[64], sm[64];
  const int p0 = blockIdx.x * 64;
  if (threadIdx.x < 64) {
    const int p = p0 + threadIdx.x;
    int a = 0, b = 0, mode = 0;
    if (p < npairs) {
      a = pairs[2 * p];
      b = pairs[2 * p + 1];
      const bool valid = a >= 0 && a < nsta && b >= 0 && b < nsta;
      mode = valid ? ((*nonfinite != 0 || dup[p]) ? 1 : 0) : (first ? 2 : 0);
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    sm[threadIdx.x] = mode;
  }
  __syncthreads();
  const size_t rows = (size_t)cw * nsta;
  const uint2* U2 = reinterpret_cast<const uint2*>(S2);
  const int j = blockIdx.y * 256 + threadIdx.x;
  for (int q = 0; q < 64; q++) {
    const int mode = sm[q];
    if (mode == 0 || j >= nf) continue;
    const int p = p0 + q, a = sa[q], b = sb[q];
    float re = 0.0f, im = 0.0f;
    if (mode == 1) {
      for (int wl = 0; wl < cw; wl++) {
        const unsigned char* u = use + (size_t)(w0 + wl) * nsta;
        if (!(u[a] && u[b])) continue;   // uniform over the block
        const size_t ka = (size_t)j * rows + (size_t)wl * nsta + a;
        const size_t kb = (size_t)j * rows + (size_t)wl * nsta + b;
        const float2 xa = SPLIT ? join16(U2[ka]) : S2[ka];
        const float2 xb = SPLIT ? join16(U2[kb]) : S2[kb];
        re = __builtin_fmaf(xa.x, xb.x, __builtin_fmaf(xa.y, xb.y, re));
        im = __builtin_fmaf(xa.x, xb.y, __builtin_fmaf(-xa.y, xb.x, im));
      }
    }
    float2* o = acc + (size_t)p * nf + j;
    float2 v = first ? make_float2(0.0f, 0.0f) : *o;
    v.x += re;
    v.y += im;
    *o = v;
  }
}

__global__ void k_cc_fill_i32(int* __restrict__ x, size_t n, int v) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = v;
}

// acc[k] = (re[k], im[k]): running sums from separate arrays
__global__ void k_cc_interleave(const float* __restrict__ re, const float* __restrict__ im,
                                size_t n, float2* __restrict__ acc) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) acc[k] = make_float2(re[k], im[k]);
}

// ncfsr/ncfsi[p][j] = acc[p][j] / count[p] where count > 0, else acc
// (the untouched running sums, as the reference leaves them)
// With im == nullptr, re holds interleaved complex [npairs][nf][2] instead.
// re/im[k] (or interleaved re[2k], re[2k+1]) = acc[k] / count[k / nf]
// (acc[k] itself when the count is 0), k over the flat [npairs][nf]; V = 4
// consecutive values a thread with 16-byte accesses (n a multiple of 4,
// the outputs 16-byte aligned), else one.  A block per (pair, 256 bins)
// made 80k blocks for 19,900 pairs x 1000 bins (0.071 ms; flat, 0.064).
template <int V>
__global__ __launch_bounds__(256) void k_cc_finish(const float2* __restrict__ acc,
                                                   const float* __restrict__ count, int nf,
                                                   size_t n, float* __restrict__ re,
                                                   float* __restrict__ im) {
  const size_t k0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * V;
  if (k0 >= n) return;
  float a[V], b[V];
#pragma unroll
  for (int u = 0; u < V; u++) {
    const size_t k = k0 + u;
    const float c = count[k / (size_t)nf];
    const float2 v = acc[k];
    a[u] = c > 0.0f ? v.x / c : v.x;
    b[u] = c > 0.0f ? v.y / c : v.y;
  }
  if (V == 4) {
    if (im) {
      *reinterpret_cast<float4*>(re + k0) = make_float4(a[0], a[1], a[2], a[3]);
      *reinterpret_cast<float4*>(im + k0) = make_float4(b[0], b[1], b[2], b[3]);
    } else {
      float4* o = reinterpret_cast<float4*>(re + 2 * k0);
      o[0] = make_float4(a[0], b[0], a[1], b[1]);
      o[1] = make_float4(a[2], b[2], a[3], b[3]);
    }
  } else if (im) {
    re[k0] = a[0];
    im[k0] = b[0];
  } else {
    reinterpret_cast<float2*>(re)[k0] = make_float2(a[0], b[0]);
  }
}

hipError_t launch_cc_onebit(float* x, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cc_onebit, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n);
  return hipGetLastError();
}

hipError_t launch_cc_gather(const float* data, int nsta, int npts, int L, int fold, int steplen,
                            int w0, int cw, float* seg, hipStream_t st) {
  const long long rows = (long long)cw * nsta, per = (rows + 7) / 8;
  if (rows == 0) return hipSuccess;
  if (8 * per > 0x7fffffffLL) return hipErrorInvalidValue;
  // V = 4 needs 16-byte rows of seg and an array that ends on the 16-byte
  // grid (the shifted loads read whole aligned 16 bytes)
  if (L % 4 == 0 && (size_t)nsta * npts % 4 == 0 && ((uintptr_t)data & 15) == 0 &&
      ((uintptr_t)seg & 15) == 0)
    hipLaunchKernelGGL(k_cc_gather<4>, dim3((unsigned)(8 * per)), dim3(256), 0, st, data, npts, L,
                       fold, steplen, w0, cw, nsta, seg);
  else
    hipLaunchKernelGGL(k_cc_gather<1>, dim3((unsigned)(8 * per)), dim3(256), 0, st, data, npts, L,
                       fold, steplen, w0, cw, nsta, seg);
  return hipGetLastError();
}

hipError_t launch_cc_spec(const void* spec, int rows, int nbins, int zhalf, int nf, int fstride,
                          int whiten, void* S, hipStream_t st) {
  hipLaunchKernelGGL(k_cc_spec, dim3(rows, (nf + 255) / 256), dim3(256), 0, st,
                     (const float2*)spec, nbins, zhalf, nf, fstride, whiten, (float2*)S);
  return hipGetLastError();
}

hipError_t launch_cc_use(const int* startend, int nsta, int nshifts, int steplen, int fftlen,
                         unsigned char* use, hipStream_t st) {
  const size_t n = (size_t)nshifts * nsta;
  hipLaunchKernelGGL(k_cc_use, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, startend, nsta,
                     n, steplen, fftlen, use);
  return hipGetLastError();
}

hipError_t launch_cc_count(const unsigned char* use, const int* pairs, int npairs, int nsta,
                           int nshifts, float* count, hipStream_t st) {
  if (npairs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cc_count, dim3((npairs + 3) / 4), dim3(256), 0, st, use, pairs, npairs,
                     nsta, nshifts, count);
  return hipGetLastError();
}

hipError_t launch_cc_pairs(const void* S, const unsigned char* use, const int* pairs, int npairs,
                           int nsta, int nf, int w0, int cw, void* acc, hipStream_t st) {
  hipLaunchKernelGGL(k_cc_pairs, dim3(npairs, (nf + 255) / 256), dim3(256), 0, st,
                     (const float2*)S, use, pairs, nsta, nf, w0, cw, (float2*)acc);
  return hipGetLastError();
}

hipError_t launch_cc_spec_t(const void* spec, int rows, int nsta, int nbins, int zhalf, int nf,
                            int fstride, int whiten, const unsigned char* use, int w0, void* S2,
                            int* nonfinite, hipStream_t st) {
  hipLaunchKernelGGL(k_cc_spec_t, dim3((rows + 31) / 32, (nf + 31) / 32), dim3(256), 0, st,
                     (const float2*)spec, rows, nsta, nbins, zhalf, nf, fstride, whiten, use, w0,
                     (float2*)S2, nonfinite);
  return hipGetLastError();
}

hipError_t launch_cc_pair_map(const int* pairs, int npairs, int nsta, int ntile, int* pidx,
                              unsigned char* dup, unsigned char* tiles, hipStream_t st) {
  hipLaunchKernelGGL(k_cc_pair_map, dim3((npairs + 255) / 256), dim3(256), 0, st, pairs, npairs,
                     nsta, ntile, pidx, dup, tiles);
  return hipGetLastError();
}

hipError_t launch_cc_gram(const void* S2, int nsta, int nf, int cw, int ntile, int whiten,
                          const unsigned char* tiles, const int* pidx, const int* nonfinite,
                          int first, void* acc, hipStream_t st) {
  const int groups = (nf + 3) / 4;                 // 4 bins per workgroup
  // 8 XCDs x runs of 4 groups
  const long long nwg = 32LL * ((groups + 31) / 32) * ntile * ntile;
  if (nwg > 0x7fffffffLL) return hipErrorInvalidValue;
  if (whiten)
    hipLaunchKernelGGL(k_cc_gram_split, dim3((unsigned)nwg), dim3(256), 0, st, (const uint2*)S2,
                       nsta, nf, cw, ntile, tiles, pidx, nonfinite, first, (float2*)acc);
  else
    hipLaunchKernelGGL(k_cc_gram, dim3((unsigned)nwg), dim3(256), 0, st, (const float2*)S2, nsta,
                       nf, cw, ntile, tiles, pidx, nonfinite, first, (float2*)acc);
  return hipGetLastError();
}

hipError_t launch_cc_pairs_t(const void* S2, int whiten, const unsigned char* use,
                             const int* pairs, int npairs, int nsta, int nf, int w0, int cw,
                             const unsigned char* dup, const int* nonfinite, int first,
                             void* acc, hipStream_t st) {
  // 64 pairs a block: most pairs have nothing to do here, and a block per
  // (pair, 256 bins) made 80k blocks for 19,900 pairs (21 us); 4096 blocks
  // striding over those items, 38 us
  if (npairs == 0 || nf == 0) return hipSuccess;
  if ((nf + 255) / 256 > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(whiten ? k_cc_pairs_t<true> : k_cc_pairs_t<false>,
                     dim3((npairs + 63) / 64, (nf + 255) / 256), dim3(256), 0, st,
                     (const float2*)S2, use, pairs, npairs, nsta, nf, w0, cw, dup, nonfinite,
                     first, (float2*)acc);
  return hipGetLastError();
}

hipError_t launch_cc_fill_i32(int* x, size_t n, int v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cc_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n, v);
  return hipGetLastError();
}

hipError_t launch_cc_interleave(const float* re, const float* im, size_t n, void* acc,
                                hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cc_interleave, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, re, im,
                     n, (float2*)acc);
  return hipGetLastError();
}

hipError_t launch_cc_finish(const void* acc, const float* count, int npairs, int nf, float* re,
                            float* im, hipStream_t st) {
  const size_t n = (size_t)npairs * nf;
  if (n == 0) return hipSuccess;
  const bool v4 = n % 4 == 0 && ((uintptr_t)re & 15) == 0 && ((uintptr_t)im & 15) == 0 &&
                  ((uintptr_t)acc & 15) == 0;
  const size_t threads = v4 ? n / 4 : n;
  const size_t nb = (threads + 255) / 256;
  if (nb > 0x7fffffffULL) return hipErrorInvalidValue;
  if (v4)
    hipLaunchKernelGGL(k_cc_finish<4>, dim3((unsigned)nb), dim3(256), 0, st, (const float2*)acc,
                       count, nf, n, re, im);
  else
    hipLaunchKernelGGL(k_cc_finish<1>, dim3((unsigned)nb), dim3(256), 0, st, (const float2*)acc,
                       count, nf, n, re, im);
  return hipGetLastError();
}

}  // namespace fjk
