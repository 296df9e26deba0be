// fj_mwfj.hip -- multi-window F-J (MWFJ) on the device.
//
// Reference: MWFJ (src/ccfj.pyx:252-268) builds, for every window w and
// trace j, win(npts, Fs, winl[w,j], winr[w,j], taper) (src/ccfj.pyx:148-158),
// multiplies, and calls fj_earthquake per window (src/ccfj.pyx:229-250):
// rFFT over time, then |FJ[Re] + i FJ[Im]| (func 0) or the Hankel form
// (func 1), normalised per frequency column, stored as float64.
//
// Here every stage runs on the GPU:
//   k_window        tapered traces in double (window from its integer
//                   description n1, n2, delta; degenerate windows are passed
//                   as explicit rows), a chunk of windows at a time
//   hipFFT          the rFFT, double precision like numpy's: for an even
//                   trace length a length-npts/2 complex FFT (Z2Z) of the
//                   packed samples, the used bins formed from it on read
//                   (rfft_bin); else D2Z
//   k_mwfj_coef     bins -> float32 spectra (the reference's cast) ->
//                   trapezoid coefficients B[f][k][rhs], rhs = (window,
//                   Re/Im); for func 1 the K axis holds a J half and a Y half
//   k_mwfj_mfma     itype 0: the windows become right-hand sides of one
//                   contraction per frequency, O[c][rhs] = sum_k Z(k_c r_k)
//                   B[k][rhs]; the kernel matrix Z (J0 or Y0) is generated by
//                   VALU and contracted on the matrix cores, in fp16 split
//                   products (v_mfma_f32_32x32x16_f16) or in FP32
//                   (v_mfma_f32_32x32x2_f32), 32 velocities x 128
//                   right-hand sides per wave
//   finish          |Re + i Im| was formed in the MFMA epilogue; per-column
//                   max (k_rowmax), divide and transpose to [nwin][nc][nf]
//                   in float64 through LDS tiles (k_finish<double>)
// itype 1 reuses the single-array kernels per window on the device spectra.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "fj_lane.h"
#include "fj_kernels.h"
#include "fj_mwfj.h"

namespace fjk {

using namespace fjl;

typedef float floatx16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// Taper value of window description desc at sample t.  A description whose
// xidx is outside [-1, nx) gives a zero row (the device entry point cannot
// check descriptions on the host).
__device__ inline double win_value(const int* __restrict__ n1, const int* __restrict__ n2,
                                   const int* __restrict__ dl, const int* __restrict__ xidx,
                                   const double* __restrict__ xrows, int nx, int npts, int desc,
                                   int t) {
  const int xi = xidx[desc];
  if (xi >= nx || xi < -1) return 0.0;
  if (xi >= 0) return xrows[(size_t)xi * npts + t];
  const int a = n1[desc], b = n2[desc], d = dl[desc];
  if (t < a || t >= b) return 0.0;
  if (t < a + d) return sin(3.141592653589793 / 2 * (double)(t - a) / (double)d);
  if (t >= b - d) return sin(3.141592653589793 / 2 * (double)(b - 1 - t) / (double)d);
  return 1.0;
}

// tapered traces: xw[(wl*nr + s)*npts + t], windows w0 .. w0+cw-1.  Each
// thread reads its PER consecutive samples of u[s] once and writes them,
// tapered, for every window of the chunk (the trace is read once per chunk,
// not per window).  PER = 2 (16-byte stores; npts even, so rows stay
// 16-byte aligned) or 1.  The kernel is bound by its stores (the tapered
// traces of a chunk of windows, zeros included).
template <int PER>
__global__ void k_window(const double* __restrict__ u, const int* __restrict__ n1,
                         const int* __restrict__ n2, const int* __restrict__ dl,
                         const int* __restrict__ xidx, const double* __restrict__ xrows,
                         int nx, int w0, int cw, int nr, int npts, double* __restrict__ xw) {
  const int t = PER * (blockIdx.x * blockDim.x + threadIdx.x);
  const int s = blockIdx.y;
  if (t >= npts) return;
  double us[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) us[q] = u[(size_t)s * npts + t + q];
  for (int wl = 0; wl < cw; wl++) {
    const int desc = (w0 + wl) * nr + s;   // uniform over the block
    double* dst = xw + ((size_t)wl * nr + s) * npts + t;
    if (PER == 2) {
      const double2 v = make_double2(us[0] * win_value(n1, n2, dl, xidx, xrows, nx, npts, desc, t),
                                     us[1] * win_value(n1, n2, dl, xidx, xrows, nx, npts, desc, t + 1));
      *reinterpret_cast<double2*>(dst) = v;
    } else {
      dst[0] = us[0] * win_value(n1, n2, dl, xidx, xrows, nx, npts, desc, t);
    }
  }
}

// Used bins of the rFFT of the tapered traces by the Goertzel recurrence
// over each window's support only: spec[(wl nr + s) nf + f] = X[f fstride]
// for window w0 + wl and station s, where X[k] = sum_t x[t] e^{-2 pi i k t /
// npts} and x = u[s] times the window, zero outside [a, b) = [n1, n2) (the
// whole row for an explicit row, empty for an invalid description).  With
// w = 2 pi k / npts and s_t = x_t + 2 cos(w) s_{t-1} - s_{t-2} over t = a ..
// b-1 from s = 0, sum_t x_t e^{-i w t} = e^{-i w (b-1)} (s_{b-1} - e^{-i w}
// s_{b-2}).  Two double-precision operations per sample and bin, against
// writing and transforming the whole zero-padded row: the group-velocity
// windows of an active-source gather cover about 1% of it.  The phase
// factor's argument is reduced exactly in integers (k (b-1) mod npts).
// One workgroup per (window, station); the support is staged through LDS in
// segments, 4 bins per thread per pass over it.  tw[m] = (cos, sin) of
// 2 pi m / npts (k_mwfj_twiddles) gives every trigonometric value.  A trace with a non-finite
// sample anywhere gives NaN in every bin of every window, as the transform
// of the whole tapered row does (NaN times a zero taper is NaN).
constexpr int kGzSeg = 2048;
constexpr int kGzBins = 4;
__global__ __launch_bounds__(256) void k_mwfj_goertzel(
    const double* __restrict__ u, const int* __restrict__ n1, const int* __restrict__ n2,
    const int* __restrict__ dl, const int* __restrict__ xidx, const double* __restrict__ xrows,
    int nx, int w0, int nr, int npts, int nf, int fstride, const int* __restrict__ rowbad,
    const double2* __restrict__ tw, double2* __restrict__ spec) {
  __shared__ __attribute__((aligned(16))) double xs[kGzSeg];
  const int s = blockIdx.x, wl = blockIdx.y;
  if (rowbad[s]) {
    double2* __restrict__ o = spec + ((size_t)wl * nr + s) * nf;
    for (int f = threadIdx.x; f < nf; f += 256) o[f] = make_double2(__builtin_nan(""), __builtin_nan(""));
    return;
  }
  const int desc = (w0 + wl) * nr + s;
  const int xi = xidx[desc];
  int a = 0, b = 0;
  if (xi >= 0 && xi < nx) {
    b = npts;
  } else if (xi == -1) {
    a = max(n1[desc], 0);
    b = min(n2[desc], npts);
  }
  if (b < a) b = a;
  double2* __restrict__ out = spec + ((size_t)wl * nr + s) * nf;
  const double* __restrict__ us = u + (size_t)s * npts;
  for (int fb = 0; fb < nf; fb += 256 * kGzBins) {
    double c2[kGzBins], s1[kGzBins], s2[kGzBins];
    int kk[kGzBins];
#pragma unroll
    for (int i = 0; i < kGzBins; i++) {
      const int f = min(fb + (int)threadIdx.x + 256 * i, nf - 1);
      kk[i] = f * fstride;
      c2[i] = 2.0 * tw[kk[i]].x;
      s1[i] = s2[i] = 0.0;
    }
    for (int t0 = a; t0 < b; t0 += kGzSeg) {
      const int n = min(kGzSeg, b - t0);
      __syncthreads();   // the previous segment has been read
      for (int q = threadIdx.x; q < n; q += 256)
        xs[q] = us[t0 + q] * win_value(n1, n2, dl, xidx, xrows, nx, npts, desc, t0 + q);
      __syncthreads();
      // two samples a step, the recurrence's two values taking turns in
      // place (no register moves): after x0, s2 holds the newest value;
      // after x1, s1 does and s2 the one before it
      // (the next pair is read from LDS while this one is used)
      int q = 0;
      const double2* xp = reinterpret_cast<const double2*>(xs);
      double2 nxt = n > 1 ? xp[0] : make_double2(0.0, 0.0);
      for (; q + 1 < n; q += 2) {
        const double2 cur = nxt;
        if (q + 3 < n) nxt = xp[(q >> 1) + 1];
#pragma unroll
        for (int i = 0; i < kGzBins; i++) {
          s2[i] = __builtin_fma(c2[i], s1[i], cur.x - s2[i]);
          s1[i] = __builtin_fma(c2[i], s2[i], cur.y - s1[i]);
        }
      }
      if (q < n) {
        const double x = xs[q];
#pragma unroll
        for (int i = 0; i < kGzBins; i++) {
          const double v = __builtin_fma(c2[i], s1[i], x - s2[i]);
          s2[i] = s1[i];
          s1[i] = v;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kGzBins; i++) {
      const int f = fb + (int)threadIdx.x + 256 * i;
      if (f >= nf) continue;
      double2 X = make_double2(0.0, 0.0);
      if (b > a) {
        const int k = kk[i];
        const double2 e1 = tw[k];   // e^{-i w} = e1.x - i e1.y
        const double yr = s1[i] - e1.x * s2[i], yi = e1.y * s2[i];
        const long long m = ((long long)k * (long long)(b - 1)) % npts;
        const double2 em = tw[m];   // e^{-i w (b-1)} = em.x - i em.y
        X = make_double2(em.x * yr + em.y * yi, em.x * yi - em.y * yr);
      }
      out[f] = X;
    }
  }
}

// tw[m] = (cos, sin)(2 pi m / n), m < n
__global__ void k_mwfj_twiddles(int n, double2* __restrict__ tw) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double sn, cs;
  sincospi(2.0 * (double)m / (double)n, &sn, &cs);
  tw[m] = make_double2(cs, sn);
}

// rowbad[s] = 1 when trace s has a non-finite sample (k_mwfj_goertzel)
__global__ __launch_bounds__(256) void k_mwfj_rowbad(const double* __restrict__ u, int npts,
                                                     int* __restrict__ rowbad) {
  const int s = blockIdx.x;
  const double* row = u + (size_t)s * npts;
  int any = 0;   // v - v != 0: NaN or inf
  if ((npts & 1) == 0 && ((uintptr_t)u & 15) == 0) {   // uniform: 16-byte loads
    const double2* r2 = reinterpret_cast<const double2*>(row);
#pragma unroll 8
    for (int t = threadIdx.x; t < npts / 2; t += 256) {
      const double2 v = r2[t];
      any |= !(v.x - v.x == 0.0) | !(v.y - v.y == 0.0);
    }
  } else {
#pragma unroll 8
    for (int t = threadIdx.x; t < npts; t += 256) any |= !(row[t] - row[t] == 0.0);
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) rowbad[s] = any;
}

// Sum over window descriptions of the support length b - a (as in
// k_mwfj_goertzel), for the choice between it and the FFT.
__global__ void k_mwfj_support(const int* __restrict__ n1, const int* __restrict__ n2,
                               const int* __restrict__ xidx, int nx, int npts, int ndesc,
                               unsigned long long* __restrict__ sum) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (d < ndesc) {
    const int xi = xidx[d];
    if (xi >= 0 && xi < nx) {
      v = (unsigned long long)npts;
    } else if (xi == -1) {
      const int a = max(n1[d], 0), b = min(n2[d], npts);
      v = b > a ? (unsigned long long)(b - a) : 0ull;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(sum, v);
}

// Bin k of the real DFT of a row of npts = 2M samples.  zhalf = false: row
// holds the rFFT itself (M + 1 bins).  zhalf = true: row holds the length-M
// complex DFT Z of z[n] = x[2n] + i x[2n+1], and for k <= M
//   X[k] = E[k] + e^{-i pi k / M} O[k],
//   E[k] = (Z[k] + conj Z[M-k]) / 2,  O[k] = (Z[k] - conj Z[M-k]) / (2i),
// with Z[M] = Z[0]: only the bins that are used are formed.
__device__ inline double2 rfft_bin(const double2* __restrict__ row, int k, int M, bool zhalf) {
  if (!zhalf) return row[k];
  const double2 a = row[k < M ? k : 0];
  const double2 b = row[k == 0 ? 0 : M - k];   // Z[M-k], Z[M] = Z[0]
  const double er = 0.5 * (a.x + b.x), ei = 0.5 * (a.y - b.y);
  const double orr = 0.5 * (a.y + b.y), oi = -0.5 * (a.x - b.x);
  double sn, cs;
  sincospi((double)k / (double)M, &sn, &cs);   // e^{-i pi k/M} = cs - i sn
  return make_double2(er + cs * orr + sn * oi, ei + cs * oi - sn * orr);
}

// spectra -> float32 staging -> trapezoid coefficients.
// B layout [nf][Kp][nrhs]; rhs 2w = Re, 2w+1 = Im of window w.
// func 0: rows k = s.  func 1: rows s (J half: Re, Im) and half + s
// (Y half: -Im, Re), so O_re = FJ[Re] - FY[Im], O_im = FJ[Im] + FY[Re].
// A block takes one station, 64 frequencies and 64 windows of the chunk:
// it reads the spectra along the bins (coalesced; one twiddle per bin for
// all 64 windows), transposes through LDS and writes each B row segment of
// 64 windows x (Re, Im) = 512 bytes in one piece.
constexpr int kCoefT = 64;
__global__ __launch_bounds__(256) void k_mwfj_coef(const double2* __restrict__ spec,
                                                   const float* __restrict__ r, int nr, int w0,
                                                   int cw, int nf, int fstride, int nbins,
                                                   int zhalf, int func, int half, int nrhs,
                                                   float* __restrict__ B,
                                                   unsigned int* __restrict__ bmax) {
  __shared__ float tre[kCoefT][kCoefT + 1], tim[kCoefT][kCoefT + 1];
  __shared__ double tsn[kCoefT], tcs[kCoefT];
  const int f0 = blockIdx.x * kCoefT;
  const int s = blockIdx.y;
  const int wb = blockIdx.z * kCoefT;
  const int M = nbins;
  if (threadIdx.x < kCoefT) {
    const int k = (f0 + threadIdx.x) * fstride;
    double sn = 0.0, cs = 1.0;
    if (zhalf) sincospi((double)k / (double)M, &sn, &cs);   // e^{-i pi k/M} = cs - i sn
    tsn[threadIdx.x] = sn;
    tcs[threadIdx.x] = cs;
  }
  const float rs = r[s];
  double w = 0.0;
  if (s > 0) w += (double)fmaxf(rs - r[s - 1], 0.1f);
  if (s + 1 < nr) w += (double)fmaxf(r[s + 1] - rs, 0.1f);
  // trapezoid weight x r, over sqrt(r): the kernel matrix carries
  // sqrt(r) Z0(k r) (its large-argument amplitude without 1/sqrt r) and the
  // contraction's epilogue multiplies by sqrt(2/(pi k))
  const double wr = 0.5 * w * sqrt((double)rs);
  __syncthreads();
  for (int it = threadIdx.x; it < kCoefT * kCoefT; it += 256) {
    const int fl = it & (kCoefT - 1), wl = it / kCoefT;
    const int f = f0 + fl, wc = wb + wl;
    float are = 0.0f, aim = 0.0f;
    if (f < nf && wc < cw) {
      const double2* row = spec + ((size_t)wc * nr + s) * nbins;
      const int k = f * fstride;
      double zr, zi;
      if (!zhalf) {
        const double2 z = row[k];
        zr = z.x;
        zi = z.y;
      } else {   // X[k] = E[k] + e^{-i pi k/M} O[k] (rfft_bin)
        const double2 a = row[k < M ? k : 0];
        const double2 b = row[k == 0 ? 0 : M - k];   // Z[M-k], Z[M] = Z[0]
        const double er = 0.5 * (a.x + b.x), ei = 0.5 * (a.y - b.y);
        const double orr = 0.5 * (a.y + b.y), oi = -0.5 * (a.x - b.x);
        const double cs = tcs[fl], sn = tsn[fl];
        zr = er + cs * orr + sn * oi;
        zi = ei + cs * oi - sn * orr;
      }
      const float re = (float)zr, im = (float)zi;   // the reference's float32 staging
      are = (float)(wr * (double)re);
      aim = (float)(wr * (double)im);
    }
    tre[fl][wl] = are;
    tim[fl][wl] = aim;
  }
  __syncthreads();
  // max |B| of each frequency, as float bits (inf and NaN order above every
  // finite value), for the contraction's choice of form and scale
  if (threadIdx.x < kCoefT && f0 + (int)threadIdx.x < nf) {
    unsigned int m = 0u;
    for (int wl = 0; wl < kCoefT; wl++) {
      m = max(m, __float_as_uint(fabsf(tre[threadIdx.x][wl])));
      m = max(m, __float_as_uint(fabsf(tim[threadIdx.x][wl])));
    }
    if (m) atomicMax(&bmax[f0 + threadIdx.x], m);
  }
  const int Kp = func ? 2 * half : half;
  for (int it = threadIdx.x; it < kCoefT * 2 * kCoefT; it += 256) {
    const int fl = it / (2 * kCoefT), e = it & (2 * kCoefT - 1);
    const int wl = e >> 1, part = e & 1;
    const int f = f0 + fl, wc = wb + wl;
    if (f >= nf || wc >= cw) continue;
    const int col = 2 * (w0 + wc) + part;
    const float are = tre[fl][wl], aim = tim[fl][wl];
    B[((size_t)f * Kp + s) * nrhs + col] = part ? aim : are;
    if (func) B[((size_t)f * Kp + half + s) * nrhs + col] = part ? are : -aim;
  }
}

// padded station arrays: rr (double) and ir = 1/r (float), ntot >= nr
// entries, padding repeating the last distance (its rows of B are zero)
__global__ void k_mwfj_stations(const float* __restrict__ r, int nr, int ntot,
                                double* __restrict__ rr, float* __restrict__ ir) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ntot) return;
  const double v = (double)r[s < nr ? s : nr - 1];
  rr[s] = v;
  ir[s] = (float)(1.0 / v);
}

// sqrt(2/(pi k)) * (that value) = Z0(k r) sqrt(r): the kernel-matrix entry
// of one lane (velocity) and station.  x >= kFarX: the modulus-phase form
// (one transcendental; in a FAR block, where every lane of the wave has
// x >= kFarX, with no branch); otherwise the large-argument fits for x >= 8
// and the double polynomials below.  Both paths give an entry the same
// value, whichever lanes share its wave.
template <bool Y, bool FAR>
__device__ inline float mwfj_value(double rs, float ir, const LaneK& L) {
  const float z = L.zk * ir;   // 8 / x
  if (FAR || z <= 8.0f / kFarX) {
    const float t = z * z;
    const float u = __builtin_fmaf(z, fitf_ph0(t), phase_rev(L.q, rs));
    return fitf_M0(t) * (Y ? sin_rev(u) : cos_rev(u));
  }
  if (z <= 1.0f) {
    const float t = z * z;
    const float p0 = fit_P0(t);
    const float q0 = z * fit_q0(t);
    float sn, co;
    sincos_rev(phase_rev(L.q, rs), &sn, &co);
    return Y ? __builtin_fmaf(p0, sn, q0 * co) : __builtin_fmaf(p0, co, -q0 * sn);
  }
  const double xd = L.q2pi * rs;
  double y, Lg;
  const Bess b = bessel_small<false, Y, FIT_TABLE>(xd, &y, &Lg);
  const float ia = L.kf > 0.0f ? sqrtf(kPiOver2 * (float)xd) : sqrtf((float)rs);
  return (Y ? b.y0 : b.j0) * ia;
}

// O[c][rhs] = sum_k Z(k_c r_k) B[k][rhs] for one frequency, 128 velocities
// (4 waves x 32) and 128 right-hand sides (4 MFMA tiles of 32) per workgroup,
// K in chunks of 32 stations double-buffered in LDS: chunk c+1 is fetched
// into registers while chunk c is contracted (one barrier per chunk).  Each
// wave first computes its lane's 16 kernel-matrix values of the chunk as
// independent chains, then issues the chunk's MFMAs back to back, so one
// wave's Bessel work overlaps another's matrix work.
//
// Two forms, chosen per workgroup (one frequency) from bmax[f], the bits of
// max |B| over the frequency's coefficients (atomicMax in k_mwfj_coef; the
// bits of inf and NaN order above every finite value):
//   split (f > 0 and every coefficient finite): each FP32 operand held as
//     two fp16 numbers, v = v_hi + v_lo (v_hi = v rounded to fp16, v_lo the
//     rounded remainder; 22 significand bits together), each product as
//     three fp16 products a_hi b_hi + a_hi b_lo + a_lo b_hi accumulated in
//     FP32 by v_mfma_f32_32x32x16_f16; the dropped a_lo b_lo and the
//     rounding of the low parts leave about 2^-21 of each product.  The
//     entries sqrt(r) Z0(k r) are at most about 2.5 in size for k > 0; B is
//     scaled by the power of two 2^(13 - floor(log2 max |B|)) (exact), so
//     that |B| < 2^14.  A 32x32x16 fp16 MFMA takes 16 stations in 32 cycles
//     against 2 in 64 for the FP32 form: the three products need 5.3x less
//     matrix time.
//   FP32 (f = 0, where the entries are sqrt(r) Z0(0), unbounded in r and
//     -inf for Y; or a non-finite coefficient, whose NaN pattern this form
//     keeps): v_mfma_f32_32x32x2_f32 on the FP32 values.
// Operand layouts (lane l): FP32 A[i = l%32][k = l/32], B[k = l/32][n =
// l%32]; fp16 A[i = l%32][k = 8(l/32) + q], B[k = 8(l/32) + q][n = l%32],
// q = 0..7 (A and B share the k mapping, which is all the product needs).
// Both write the result layout of every 32x32 MFMA: register v holds row
// 8(v/4) + 4(l/32) + v%4, column l%32.
constexpr int kMwC = 128;   // velocities per workgroup
constexpr int kMwR = 128;   // right-hand sides per workgroup
constexpr int kMwK = 32;    // K rows per LDS chunk
constexpr int kHP = 40;     // row pitch of the split form's transposed B planes, in
                            // halves (80 bytes: the 8 lanes of a 16-byte read phase
                            // hit distinct banks)
// static LDS of a workgroup: the FP32 form's [2][kMwK][kMwR + 4] floats
// (33,792 B) or the split form's [2][2][kMwR][kHP] halves (40,960 B)
constexpr int kMwSmem = 2 * 2 * kMwR * kHP * 2;
static_assert(2 * kMwK * (kMwR + 4) * 4 <= kMwSmem, "LDS of the FP32 form");

// waves per SIMD the contraction kernel is held to at least (a tuning
// switch: -DFJ_MWFJ_WAVES=N)
#ifndef FJ_MWFJ_WAVES
#define FJ_MWFJ_WAVES 3
#endif

// Station arrays of up to kMwLdsSt rows are copied into LDS (dynamic shared
// memory, 12 bytes a row) at the start, so that the VALU phase reads them
// there instead of waiting on a global load per entry; larger arrays read
// them from global memory.
constexpr int kMwLdsSt = 1536;

typedef _Float16 half8 __attribute__((ext_vector_type(8)));

// The lane's 16 kernel-matrix values of a chunk at stations s(h), h = 0..15.
// When x >= kFarX at every one of them on every lane, they are 16
// independent straight-line chains; otherwise each station takes its own
// wave-uniform branch.  Both give an entry the same value.
template <bool Y, typename SF>
__device__ inline void mwfj_chunk_values(const double* __restrict__ rrs,
                                         const float* __restrict__ irs, const LaneK& L,
                                         SF station, float* av) {
  float irm = 0.0f;
#pragma unroll
  for (int h = 0; h < 16; h++) irm = fmaxf(irm, irs[station(h)]);
  if (!wave_any(!(L.zk * irm <= 8.0f / kFarX))) {
#pragma unroll
    for (int h = 0; h < 16; h++) av[h] = mwfj_value<Y, true>(rrs[station(h)], irs[station(h)], L);
  } else {
#pragma unroll
    for (int h = 0; h < 16; h++) {
      const int s = station(h);
      const float ir = irs[s];
      const bool far = !wave_any(L.zk * ir > 8.0f / kFarX);
      av[h] = far ? mwfj_value<Y, true>(rrs[s], ir, L) : mwfj_value<Y, false>(rrs[s], ir, L);
    }
  }
}

// |Re + i Im| per window (Re in even columns, Im in odd) times the amplitude
// sqrt(2/(pi k)) of the row's velocity (1 at k = 0) and `unscale`, into T;
// and the maximum over the wave's 32 velocities of each (window,
// frequency) into Mp[g][w][f], g = the wave's velocity group (reduced over
// g by k_mwfj_colmax).  As in k_rowmax: fmaxf from 0 skips a NaN, and
// group 0 holds |value| at velocity 0 instead when that is NaN.
__device__ inline void mwfj_epilogue(const floatx16* acc, const float* __restrict__ fv,
                                     const float* __restrict__ cv, int fi, int rg, int wave,
                                     int i, int kq, int nf, int nc, int nwin, float unscale,
                                     float* __restrict__ T, float* __restrict__ Mp) {
  float amp[16];
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int m = 8 * (v >> 2) + 4 * kq + (v & 3);
    const int c = blockIdx.x * kMwC + wave * 32 + m;
    const double k = (double)fv[fi] / (double)cv[c < nc ? c : nc - 1] * kTwoPi;
    amp[v] = ((float)k > 0.0f ? (float)sqrt(2.0 / (kPiD * k)) : 1.0f) * unscale;
  }
  const bool has_c0 = blockIdx.x == 0 && wave == 0 && kq == 0;   // c = 0 at v = 0
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int col = rg * kMwR + t * 32 + i;
    float vmax = 0.0f, a0 = 0.0f;
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const float mine = acc[t][v];
      const float other = __shfl_xor(mine, 1);
      const int m = 8 * (v >> 2) + 4 * kq + (v & 3);
      const int c = blockIdx.x * kMwC + wave * 32 + m;
      const float val = sqrtf(mine * mine + other * other) * amp[v];
      if (v == 0) a0 = val;
      if (c < nc) vmax = fmaxf(vmax, val);
      if ((i & 1) == 0 && c < nc && col < 2 * nwin) {
        const int w = col >> 1;
        T[((size_t)w * nf + fi) * nc + c] = val;
      }
    }
    vmax = fmaxf(vmax, __shfl_xor(vmax, 32));   // the other 16 velocities
    if (has_c0 && a0 != a0) vmax = fabsf(a0);
    if ((i & 1) == 0 && kq == 0 && col < 2 * nwin) {
      const size_t g = (size_t)blockIdx.x * 4 + wave;
      Mp[(g * nwin + (col >> 1)) * nf + fi] = vmax;
    }
  }
}

// M[k] = the column maximum of (window, frequency) k from the velocity
// groups' partial maxima Mp[g][k]: group 0's value when NaN (a NaN at
// velocity 0), else the fmaxf of all groups' (which skip NaNs).  Equal to
// k_rowmax over T.  Writing the partials replaced a pass of k_rowmax over
// T (0.14 ms for 64 windows x 820 f x 2000 c); atomicMax on the bits from
// every wave cost the contraction 0.07 ms more.
__global__ __launch_bounds__(256) void k_mwfj_colmax(const float* __restrict__ Mp, int ng,
                                                     size_t n, float* __restrict__ M) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float v = Mp[k];
  if (v == v) {
    // eight independent chains (max is exact and order-free; groups past 0
    // hold no NaN): one chain of 63 dependent loads took 27 us
    float m[8];
#pragma unroll
    for (int u = 0; u < 8; u++) m[u] = v;
    int g = 1;
    for (; g + 8 <= ng; g += 8)
#pragma unroll
      for (int u = 0; u < 8; u++) m[u] = fmaxf(m[u], Mp[(size_t)(g + u) * n + k]);
    for (; g < ng; g++) m[0] = fmaxf(m[0], Mp[(size_t)g * n + k]);
#pragma unroll
    for (int u = 1; u < 8; u++) m[0] = fmaxf(m[0], m[u]);
    v = m[0];
  }
  M[k] = v;
}

template <int FUNC, bool LDS_ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FJ_MWFJ_WAVES, 8))) void k_mwfj_mfma(
    const float* __restrict__ B, const unsigned int* __restrict__ bmax,
    const double* __restrict__ rr, const float* __restrict__ irr, const float* __restrict__ fv,
    const float* __restrict__ cv, int nr, int half, int nf, int nc, int nrhs, int nwin,
    float* __restrict__ T, float* __restrict__ Mp) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[kMwSmem];
  extern __shared__ double st_lds[];   // [half] r, then [half] 1/r (LDS_ST)
  const int fi = blockIdx.y;
  const int rg = blockIdx.z;   // right-hand-side group
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int i = lane & 31, kq = lane >> 5;
  const int crow = blockIdx.x * kMwC + wave * 32 + i;
  const int cidx = crow < nc ? crow : nc - 1;
  const LaneK L = lane_consts((double)fv[fi] / (double)cv[cidx]);
  const int Kp = FUNC ? 2 * half : half;   // a multiple of kMwK
  const float* Bf = B + (size_t)fi * Kp * nrhs + (size_t)rg * kMwR;
  const int rcols = min(kMwR, nrhs - rg * kMwR);   // multiple of 32
  if (LDS_ST) {
    float* irl = reinterpret_cast<float*>(st_lds + half);
    for (int q = threadIdx.x; q < half; q += 256) {
      st_lds[q] = rr[q];
      irl[q] = irr[q];
    }
  }
  const double* rrs = LDS_ST ? st_lds : rr;
  const float* irs = LDS_ST ? reinterpret_cast<const float*>(st_lds + half) : irr;
  floatx16 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int v = 0; v < 16; v++) acc[t][v] = 0.0f;
  const unsigned int mb = bmax[fi];
  const bool split = mb < 0x7f800000u && fv[fi] > 0.0f;   // uniform over the block

  if (split) {
    auto Bs = reinterpret_cast<_Float16 (*)[2][kMwR][kHP]>(smem);   // [buf][hi/lo][rhs][k]
    const float sc = mb == 0u ? 1.0f : ldexpf(1.0f, 13 - ilogbf(__uint_as_float(mb)));
    // staging: thread -> column cc of the chunk, rows 16 kg .. 16 kg + 15
    const int cc = threadIdx.x & (kMwR - 1), kg = threadIdx.x >> 7;
    float nb[16];
    auto fetch = [&](int k0) {
#pragma unroll
      for (int j = 0; j < 16; j++)
        nb[j] = cc < rcols ? Bf[(size_t)(k0 + 16 * kg + j) * nrhs + cc] : 0.0f;
    };
    auto stash = [&](int buf) {
#pragma unroll
      for (int j0 = 0; j0 < 16; j0 += 8) {
        half8 hi, lo;
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const float v = nb[j0 + q] * sc;
          const _Float16 h = (_Float16)v;
          hi[q] = h;
          lo[q] = (_Float16)(v - (float)h);
        }
        *reinterpret_cast<half8*>(&Bs[buf][0][cc][16 * kg + j0]) = hi;
        *reinterpret_cast<half8*>(&Bs[buf][1][cc][16 * kg + j0]) = lo;
      }
    };
    fetch(0);
    stash(0);
    __syncthreads();
    for (int k0 = 0, buf = 0; k0 < Kp; k0 += kMwK, buf ^= 1) {
      const bool more = k0 + kMwK < Kp;
      if (more) fetch(k0 + kMwK);
      const bool ypart = FUNC && k0 >= half;
      const int sbase = ypart ? k0 - half : k0;
      // value h = 8p + q at station 16p + 8kq + q (K-half p); < half: padded arrays
      auto station = [&](int h) { return sbase + 16 * (h >> 3) + 8 * kq + (h & 7); };
      float av[16];
      if (ypart)
        mwfj_chunk_values<true>(rrs, irs, L, station, av);
      else
        mwfj_chunk_values<false>(rrs, irs, L, station, av);
      half8 ah[2], al[2];
#pragma unroll
      for (int h = 0; h < 16; h++) {
        const _Float16 x = (_Float16)av[h];
        ah[h >> 3][h & 7] = x;
        al[h >> 3][h & 7] = (_Float16)(av[h] - (float)x);
      }
      // two tiles at a time, so that consecutive MFMAs use different accumulators
#pragma unroll
      for (int p = 0; p < 2; p++)
#pragma unroll
        for (int t = 0; t < 4; t += 2) {
          const half8 bh0 = *reinterpret_cast<const half8*>(&Bs[buf][0][t * 32 + i][16 * p + 8 * kq]);
          const half8 bh1 = *reinterpret_cast<const half8*>(&Bs[buf][0][t * 32 + 32 + i][16 * p + 8 * kq]);
          const half8 bl0 = *reinterpret_cast<const half8*>(&Bs[buf][1][t * 32 + i][16 * p + 8 * kq]);
          const half8 bl1 = *reinterpret_cast<const half8*>(&Bs[buf][1][t * 32 + 32 + i][16 * p + 8 * kq]);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[p], bh0, acc[t], 0, 0, 0);
          acc[t + 1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[p], bh1, acc[t + 1], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[p], bh0, acc[t], 0, 0, 0);
          acc[t + 1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[p], bh1, acc[t + 1], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[p], bl0, acc[t], 0, 0, 0);
          acc[t + 1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[p], bl1, acc[t + 1], 0, 0, 0);
        }
      // the other buffer was last read before the previous barrier
      if (more) stash(buf ^ 1);
      __syncthreads();
    }
    mwfj_epilogue(acc, fv, cv, fi, rg, wave, i, kq, nf, nc, nwin, 1.0f / sc, T, Mp);
  } else {
    auto Bs = reinterpret_cast<float (*)[kMwK][kMwR + 4]>(smem);
    constexpr int kPer = kMwK * kMwR / 256;   // B elements per thread and chunk
    float nb[kPer];
    auto fetch = [&](int k0) {
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const int e = threadIdx.x + u * 256;
        const int kr = e / kMwR, cc = e - kr * kMwR;
        nb[u] = cc < rcols ? Bf[(size_t)(k0 + kr) * nrhs + cc] : 0.0f;
      }
    };
    auto stash = [&](int buf) {
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const int e = threadIdx.x + u * 256;
        const int kr = e / kMwR, cc = e - kr * kMwR;
        Bs[buf][kr][cc] = nb[u];
      }
    };
    fetch(0);
    stash(0);
    __syncthreads();
    for (int k0 = 0, buf = 0; k0 < Kp; k0 += kMwK, buf ^= 1) {
      const bool more = k0 + kMwK < Kp;
      if (more) fetch(k0 + kMwK);
      const bool ypart = FUNC && k0 >= half;
      const int sbase = ypart ? k0 - half : k0;
      // value h at station 2h + kq (k = l/32 of each 32x32x2 MFMA)
      auto station = [&](int h) { return sbase + 2 * h + kq; };
      float av[16];
      if (ypart)
        mwfj_chunk_values<true>(rrs, irs, L, station, av);
      else
        mwfj_chunk_values<false>(rrs, irs, L, station, av);
#pragma unroll
      for (int h = 0; h < kMwK / 2; h++) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const float b = Bs[buf][2 * h + kq][t * 32 + i];
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[h], b, acc[t], 0, 0, 0);
        }
      }
      if (more) stash(buf ^ 1);
      __syncthreads();
    }
    mwfj_epilogue(acc, fv, cv, fi, rg, wave, i, kq, nf, nc, nwin, 1.0f, T, Mp);
  }
}

// float32 staging of one window's spectra for the itype-1 path:
// re/im [nr][nf]
__global__ void k_mwfj_stage(const double2* __restrict__ spec, int nr, int wl, int nf,
                             int fstride, int nbins, int zhalf, float* __restrict__ re,
                             float* __restrict__ im) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (f >= nf) return;
  const double2 z = rfft_bin(spec + ((size_t)wl * nr + s) * nbins, f * fstride, nbins, zhalf);
  re[(size_t)s * nf + f] = (float)z.x;
  im[(size_t)s * nf + f] = (float)z.y;
}

// float32 [nc][nf] (normalised) -> float64 slice of out
__global__ void k_to_double(const float* __restrict__ a, size_t n, double* __restrict__ out) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = (double)a[k];
}

// ---------------------------------------------------------------------------
hipError_t launch_window(const MwfjDev& d, int w0, int cw, hipStream_t st) {
  if (d.npts % 2 == 0) {
    dim3 grid((d.npts / 2 + 255) / 256, d.nr);
    hipLaunchKernelGGL(k_window<2>, grid, dim3(256), 0, st, d.u, d.n1, d.n2, d.dl, d.xidx,
                       d.xrows, d.nx, w0, cw, d.nr, d.npts, d.xw);
  } else {
    dim3 grid((d.npts + 255) / 256, d.nr);
    hipLaunchKernelGGL(k_window<1>, grid, dim3(256), 0, st, d.u, d.n1, d.n2, d.dl, d.xidx,
                       d.xrows, d.nx, w0, cw, d.nr, d.npts, d.xw);
  }
  return hipGetLastError();
}

hipError_t launch_mwfj_rowbad(const MwfjDev& d, int* rowbad, hipStream_t st) {
  hipLaunchKernelGGL(k_mwfj_rowbad, dim3(d.nr), dim3(256), 0, st, d.u, d.npts, rowbad);
  return hipGetLastError();
}

hipError_t launch_mwfj_twiddles(int n, double2* tw, hipStream_t st) {
  hipLaunchKernelGGL(k_mwfj_twiddles, dim3((n + 255) / 256), dim3(256), 0, st, n, tw);
  return hipGetLastError();
}

hipError_t launch_mwfj_goertzel(const MwfjDev& d, int w0, int cw, const int* rowbad,
                                const double2* tw, double2* spec, hipStream_t st) {
  hipLaunchKernelGGL(k_mwfj_goertzel, dim3(d.nr, cw), dim3(256), 0, st, d.u, d.n1, d.n2, d.dl,
                     d.xidx, d.xrows, d.nx, w0, d.nr, d.npts, d.nf, d.fstride, rowbad, tw, spec);
  return hipGetLastError();
}

hipError_t launch_mwfj_support(const MwfjDev& d, unsigned long long* sum, hipStream_t st) {
  const int ndesc = d.nwin * d.nr;
  hipLaunchKernelGGL(k_mwfj_support, dim3((ndesc + 255) / 256), dim3(256), 0, st, d.n1, d.n2,
                     d.xidx, d.nx, d.npts, ndesc, sum);
  return hipGetLastError();
}

hipError_t launch_mwfj_stations(const MwfjDev& d, double* rr, float* ir, hipStream_t st) {
  hipLaunchKernelGGL(k_mwfj_stations, dim3((d.half + 255) / 256), dim3(256), 0, st, d.r, d.nr,
                     d.half, rr, ir);
  return hipGetLastError();
}

// Zeros in the padding of B [nf][Kp][nrhs] that k_mwfj_coef never writes:
// rows nr .. half-1 of each basis half (all columns), then columns
// 2 nwin .. nrhs-1 of rows below nr; e runs over the first set, then the
// second.  k_mwfj_coef writes every other entry on each call.
__global__ __launch_bounds__(256) void k_mwfj_bpad(float* __restrict__ B, int nf, int nh,
                                                   int half, int nr, int nrhs, int ncols,
                                                   size_t nrowpad, size_t ntotal) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < ntotal;
       e += (size_t)gridDim.x * blockDim.x) {
    size_t row, col;
    if (e < nrowpad) {   // (f, h, s in [nr, half), col in [0, nrhs))
      col = e % nrhs;
      size_t q = e / nrhs;
      const size_t s = nr + q % (size_t)(half - nr);
      q /= (size_t)(half - nr);
      const size_t h = q % nh, f = q / nh;
      row = f * (size_t)nh * half + h * half + s;
    } else {             // (f, h, s in [0, nr), col in [ncols, nrhs))
      const size_t e2 = e - nrowpad;
      col = ncols + e2 % (size_t)(nrhs - ncols);
      size_t q = e2 / (size_t)(nrhs - ncols);
      const size_t s = q % nr;
      q /= nr;
      const size_t h = q % nh, f = q / nh;
      row = f * (size_t)nh * half + h * half + s;
    }
    B[row * nrhs + col] = 0.0f;
  }
}

hipError_t launch_mwfj_bpad(const MwfjDev& d, hipStream_t st) {
  const int nh = d.func ? 2 : 1, ncols = 2 * d.nwin;
  const size_t nrowpad = (size_t)d.nf * nh * (d.half - d.nr) * d.nrhs;
  const size_t ncolpad = (size_t)d.nf * nh * d.nr * (d.nrhs - ncols);
  const size_t ntotal = nrowpad + ncolpad;
  if (ntotal == 0) return hipSuccess;   // no padding: every entry is written by k_mwfj_coef
  const unsigned nb = (unsigned)std::min<size_t>((ntotal + 255) / 256, 16384);
  hipLaunchKernelGGL(k_mwfj_bpad, dim3(nb), dim3(256), 0, st, d.B, d.nf, nh, d.half, d.nr,
                     d.nrhs, ncols, nrowpad, ntotal);
  return hipGetLastError();
}

hipError_t launch_mwfj_coef(const MwfjDev& d, int w0, int cw, hipStream_t st) {
  dim3 grid((d.nf + kCoefT - 1) / kCoefT, d.nr, (cw + kCoefT - 1) / kCoefT);
  hipLaunchKernelGGL(k_mwfj_coef, grid, dim3(256), 0, st, (const double2*)d.spec, d.r, d.nr, w0,
                     cw, d.nf, d.fstride, d.nbins, d.zhalf, d.func, d.half, d.nrhs, d.B, d.bmax);
  return hipGetLastError();
}

hipError_t launch_mwfj_mfma(const MwfjDev& d, hipStream_t st) {
  dim3 grid((d.nc + kMwC - 1) / kMwC, d.nf, (d.nrhs + kMwR - 1) / kMwR);
  const bool lds = d.half <= kMwLdsSt;
  const size_t shm = lds ? (sizeof(double) + sizeof(float)) * (size_t)d.half : 0;
  auto k = d.func ? (lds ? k_mwfj_mfma<1, true> : k_mwfj_mfma<1, false>)
                  : (lds ? k_mwfj_mfma<0, true> : k_mwfj_mfma<0, false>);
  hipLaunchKernelGGL(k, grid, dim3(256), shm, st, d.B, (const unsigned int*)d.bmax, d.rr, d.ir,
                     d.f, d.c, d.nr, d.half, d.nf, d.nc, d.nrhs, d.nwin, d.T, d.Mp);
  return hipGetLastError();
}

int mwfj_groups(int nc) { return 4 * ((nc + kMwC - 1) / kMwC); }

hipError_t launch_mwfj_finish(const MwfjDev& d, hipStream_t st) {
  const size_t n = (size_t)d.nwin * d.nf;
  if (n == 0 || d.nc <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mwfj_colmax, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d.Mp,
                     mwfj_groups(d.nc), n, d.M);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_finish_f64(d.T, d.M, d.nwin, d.nf, d.nc, d.out, false, st);
}

hipError_t launch_mwfj_stage(const MwfjDev& d, int wl, float* re, float* im, hipStream_t st) {
  dim3 grid((d.nf + 255) / 256, d.nr);
  hipLaunchKernelGGL(k_mwfj_stage, grid, dim3(256), 0, st, (const double2*)d.spec, d.nr, wl,
                     d.nf, d.fstride, d.nbins, d.zhalf, re, im);
  return hipGetLastError();
}

hipError_t launch_to_double(const float* a, size_t n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_to_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n, out);
  return hipGetLastError();
}

}  // namespace fjk
