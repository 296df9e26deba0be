// fj_lane.h -- the per-(f,c) work of the F-J kernels, shared by the gfx950
// kernels (fj_kernels.hip) and the host build of the same code
// (fj_host_emul.cpp, used by the CPU test suite to check the algorithm).
//
// What is computed (reference: src/FJcpu.cpp:230-399, the per-(f,c) sums
// behind FJ/FH, and src/ccfj.pyx:202-250 for how fj_noise / fj_earthquake
// combine them):  for every frequency f_i and velocity c_j,
//   k = 2 pi f_i / c_j,  x_s = k r_s  (stations sorted by r)
//   itype 0:  O = sum_s A_s Z0(x_s),       A_s = W_s r_s U_s
//             W_s = (d_s + d_{s+1}) / 2,   d_s = max(r_s - r_{s-1}, 0.1)
//   itype 1:  O = (1/k^2) [ sum_s alpha_s x_s Z1(x_s) - sum_s beta_s D_s ]
//             D_s = (G(x_s) - G(x_{s-1})) / (x_s - x_{s-1}),
//             G(x) = int_0^x t Z1(t) dt.
// The itype-1 form is the exact integral of the reference's linear segments
// (src/FJcpu.cpp:253-268) after one integration by parts of
// int L(rho) Z0(k rho) rho drho: rho Z0(k rho) = d/drho [rho Z1(k rho) / k].
// alpha_s is the jump of the piecewise-linear interpolant at station s --
// zero except at the two ends and beside gaps below the 0.1 m clamp -- and
// beta_s is the rise of segment s (U_s - U_{s-1}, scaled by h/0.1 when the
// clamp is active).  In this form no term is much larger than the result,
// unlike the reference's a*M0 + b*M1 closed form whose terms cancel.
// D_s for |dx| < kTau comes from the Euler-Maclaurin rule with two end
// corrections (g = x Z1, g' = x Z0, g''' = -Z1 - x Z0): no Struve function
// and no difference of nearly equal G values; its truncation error is about
// dx^6 / 30240 relative (5e-7 at the switch).
//
// Operation count.  Everything wave-uniform (per station) is precomputed
// once into the station arrays (r, gap, 1/r, 1/sqrt r) and the coefficient
// records, and arrives by scalar loads; everything per lane but
// station-independent (8/k, sqrt(2/(pi k))) is computed once per lane.
// Per station and lane that leaves, for x >= 8, z = 8/x and the amplitude
// as one multiply each, no reciprocal square root, one v_fract_f64 for the
// phase, and degree-2/3 fits beyond x = 20 (95-99% of the terms of dense
// arrays).  The trapezoid sum also moves the amplitude out of the loop:
// its coefficients carry 1/sqrt(r_s) and the lane multiplies by
// sqrt(2/(pi k)) once.
#pragma once

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/ccfj.h"
#include "fj_math.h"

// No floating-point contraction in this file: every fused multiply-add is
// written as one, and no other product and sum is fused.  A term's
// arithmetic then does not depend on how the compiler happens to shape the
// code around it (one or two velocities per lane, which kernel inlines it),
// so the item-list, batched and dense kernels agree bit for bit where they
// take the same branches.
#if defined(__clang__)
#pragma clang fp contract(off)
#endif

namespace fjl {

using namespace fjm;

constexpr int MODE_J = CCFJ_MODE_J;
constexpr int MODE_Y = CCFJ_MODE_Y;
constexpr int MODE_J_MINUS_Y = CCFJ_MODE_J_MINUS_Y;
constexpr int MODE_ABS_J = CCFJ_MODE_ABS_J;
constexpr int MODE_ABS_H = CCFJ_MODE_ABS_H;

constexpr int kPartial = 32;  // stations per float partial sum
// Station loops run in blocks of kUnroll so the wave-uniform station data of
// a block are fetched by a few wide scalar loads.  The station arrays are
// padded to a multiple of kUnroll (plus one spare block for the gap array)
// with copies of the last distance and zero coefficients, which add nothing.
constexpr int kUnroll = 8;

FJ_HD inline int padded_stations(int nr) { return (nr + kUnroll - 1) / kUnroll * kUnroll; }

// Per-station arrays of one (sub-)array, nrp + kUnroll entries each; entries
// past the last station repeat its distance with a zero gap.
struct Stations {
  const double* rr;  // r_s
  const float* hh;   // r_s - r_{s-1} (hh[0] = 0)
  const float* ir;   // 1 / r_s      (inf at r = 0, where only x < 8 code runs)
  const float* isr;  // 1 / sqrt(r_s)
  const uint32_t* bw;  // block words (station_block_word)
  const float* dr;   // r_s - r_a, a = s rounded down to a multiple of 4 (station_dr)
};

// 1/r_s and the gap left of station s as the station arrays hold them
// (entries past the last station repeat its distance, with a zero gap).
FJ_HD inline float station_ir(const float* r, int nr, int s) {
  return (float)(1.0 / (double)r[s < nr ? s : nr - 1]);
}
FJ_HD inline float station_hh(const float* r, int nr, int s) {
  return (s < nr && s > 0) ? r[s] - r[s - 1] : 0.0f;
}

// 1/sqrt(r_s) as the station arrays hold it: held to the largest float at
// r = 0, so that the linear-segment records, which carry sqrt(r_s), come
// back as zeros there (and not as 0 * inf) when multiplied by it.
FJ_HD inline float station_isr(float r) {
  const float v = (float)(1.0 / sqrt((double)r));
  return v > FLT_MAX ? FLT_MAX : v;
}

// (2/pi) a_s for the Y basis of a far station, from the hat record ah =
// a sqrt r and isr = 1/sqrt r: computed the same way by the coefficient
// preparation (the kappa plane, read by smooth blocks) and by the lanes of
// the other blocks, so both give the same float.
FJ_HD inline float kappa_of(float ah, float isr) { return (ah * isr) * kTwoOverPi; }

// Offset of station s from the anchor of its block of 4, the block's first
// station a: r_s - r_a, formed in double and rounded once (exact when r_a
// and r_s are within a factor of 2).
FJ_HD inline float station_dr(const float* r, int nr, int s) {
  const int a = s & ~3;
  return (float)((double)r[s < nr ? s : nr - 1] - (double)r[a < nr ? a : nr - 1]);
}

// Entry s of the station arrays from the sorted float distances r[0..nr).
FJ_HD inline void station_entry(const float* r, int nr, int s, double* rr, float* hh, float* ir,
                                float* isr, float* dr) {
  const float rf = r[s < nr ? s : nr - 1];
  rr[s] = (double)rf;
  hh[s] = station_hh(r, nr, s);
  ir[s] = station_ir(r, nr, s);
  isr[s] = station_isr(rf);
  dr[s] = station_dr(r, nr, s);
}

// Far argument in the F-J lanes: x >= 8, the small/large split itself.  Far
// stations take the modulus-phase (or P, Q) forms with the degree-2 fitm_
// fits on t = (8/x)^2 in (0, 1] (absolute errors at most 7.6e-8, about the
// float rounding of the terms they enter); below 8 the double-precision
// small-argument polynomials.  (The far forms once started at x = 20, with
// degree-2 fits and a third form of degree-4 fits between 8 and 20: 5% of
// the dense array's trapezoid blocks then held a station below 20 for some
// lane against 2% below 8, and such a block costs the wave every form its
// lanes need.)
constexpr float kLaneFar = kSplit;

// Per-lane constants of one (f, c).
struct LaneK {
  double q;      // f / c, cycles per metre
  float qf;      // q in float (anchored block phases, block_phases)
  double q2pi;   // k = 2 pi f / c
  float kf;      // k in float
  float zk;      // 8 / k           (z = 8/x = zk / r)
  float ak;      // sqrt(2/(pi k))   (Hankel amplitude = ak / sqrt r)
  uint32_t irf;  // bits of k / kLaneFar: x >= kLaneFar where the bits of 1/r are at most this
};

FJ_HD inline uint32_t f32_bits(float v);

FJ_HD inline LaneK lane_consts(double q) {
  LaneK L;
  L.q = q;
  L.qf = (float)q;
  L.q2pi = q * kTwoPi;
  L.kf = (float)L.q2pi;
  L.zk = (float)(8.0 / L.q2pi);
  L.ak = (float)sqrt(2.0 / (kPiD * L.q2pi));
  L.irf = f32_bits((float)(L.q2pi / (double)kLaneFar));
  return L;
}

// ---- wave-uniform far blocks -----------------------------------------------
// A block of kUnroll stations is "far" when x >= kLaneFar for every lane of
// the wave, i.e. when every station of the block has 1/r <= kmin / kLaneFar with
// kmin the smallest k of the wave.  The test is one scalar comparison per
// block (1/r are non-negative floats, so their bit patterns order as
// unsigned integers; r = 0 gives +inf, r < 0 a sign bit: never far), and a
// far block runs the far-argument code with no per-lane branch on x.  It
// needs every lane of the wave active (kernels clamp surplus lanes to a
// valid velocity and skip only their store).  The host build treats each
// lane as its own wave.
FJ_HD inline uint32_t f32_bits(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(v);
#else
  uint32_t u;
  memcpy(&u, &v, 4);
  return u;
#endif
}

FJ_HD inline uint32_t far_threshold(double q2pi) {
  double kmin = q2pi;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) kmin = fmin(kmin, __shfl_xor(kmin, o));
  return __builtin_amdgcn_readfirstlane(f32_bits((float)(kmin / (double)kLaneFar)));
#else
  return f32_bits((float)(kmin / (double)kLaneFar));
#endif
}

// Block words, one uint32 per station slot, so that a block's far, smooth
// and anchor tests are one wave-uniform load and one compare each (the bits
// of non-negative floats order as unsigned integers):
//   bw[sb]     (sb % 4 == 0)  max bits of 1/r over stations sb .. sb+4
//                             (a block of 4 and the next block's first)
//   bw[sb + 1]                max bits of the gaps hh[sb .. sb+4]
//   bw[sb + 2] (sb % 8 == 0)  max bits of 1/r over stations sb .. sb+7
//   bw[sb + 3]                max bits of |dr| over stations sb .. sb+3
//   bw[sb + 6] (sb % 8 == 0)  max bits of |dr| over stations sb .. sb+7
// other slots 0.  Built from r with the station arrays' own expressions.
// (Measured on the MI355X, the trapezoid J kernel at 19,900 pairs x 4096 f
// x 3000 c: 107.3 ms against 108.6 ms with the maximum taken per block.)
FJ_HD inline uint32_t station_block_word(const float* r, int nr, int s) {
  const int k = s & 7;
  uint32_t m = 0;
  if ((k & 3) == 0) {
    for (int i = 0; i <= 4; i++) {
      const uint32_t b = f32_bits(station_ir(r, nr, s + i));
      m = b > m ? b : m;
    }
  } else if ((k & 3) == 1) {
    for (int i = 0; i <= 4; i++) {
      const uint32_t b = f32_bits(station_hh(r, nr, s - 1 + i));
      m = b > m ? b : m;
    }
  } else if (k == 2) {
    for (int i = 0; i < 8; i++) {
      const uint32_t b = f32_bits(station_ir(r, nr, s - 2 + i));
      m = b > m ? b : m;
    }
  } else if ((k & 3) == 3 || k == 6) {
    const int n = k == 6 ? 8 : 4;
    for (int i = 0; i < n; i++) {
      const uint32_t b = f32_bits(fabsf(station_dr(r, nr, (s & ~(n - 1)) + i)));
      m = b > m ? b : m;
    }
  }
  return m;
}

// true when the predicate holds on any active lane of the wave (uniform;
// one scalar compare of the lane mask)
FJ_HD inline bool wave_any(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ballot_w64(p) != 0;
#else
  return p;
#endif
}
// A segment [x_a, x_b] uses the G difference when |dx| >= min(kTau,
// kKappa min(x_a, x_b)), else the Euler-Maclaurin rule.  kTau bounds the
// rule's truncation (dx^6/30240) for oscillating g; kKappa bounds it near
// x = 0, where g = x J1 ~ x^2 is small and g = x Y1 has a log term.  In both
// cases the G difference loses at most about 1/kKappa float ulps there.
constexpr float kTau = 0.5f;
constexpr float kKappa = 0.25f;

// Bits of the largest gap h for which no segment of the wave's lanes is long
// or takes the G difference in a far block: min(h_limit, 0.99 kTau / k_max)
// with k_max the largest |k| of the wave (the margin covers the rounding of
// dx = k h, so every lane's own test agrees).  A block whose gaps are all at
// or below it (its block word, station_block_word) is smooth.
FJ_HD inline uint32_t smooth_threshold(double q2pi, uint32_t hlim) {
  double kmax = fabs(q2pi);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) kmax = fmax(kmax, __shfl_xor(kmax, o));
#endif
  const uint32_t g = f32_bits(kmax > 0.0 ? (float)(0.99 * (double)kTau / kmax) : INFINITY);
  const uint32_t t = g < hlim ? g : hlim;
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_readfirstlane(t);
#else
  return t;
#endif
}

FJ_HD inline bool use_gdiff(float xa, float xb, float dx) {
  return dx != 0.0f && fabsf(dx) >= fminf(kTau, kKappa * fminf(xa, xb));
}

// The decision for the segment to the right of a station at x with width
// dx, taken once there and reused as the left segment of the next station.
FJ_HD inline bool use_gdiff_right(float x, float dx) { return use_gdiff(x, x + dx, dx); }

template <int MODE>
struct ModeT;
template <> struct ModeT<MODE_J> { static constexpr int nspec = 1; static constexpr bool J = true, Y = false; };
template <> struct ModeT<MODE_Y> { static constexpr int nspec = 1; static constexpr bool J = false, Y = true; };
template <> struct ModeT<MODE_J_MINUS_Y> { static constexpr int nspec = 2; static constexpr bool J = true, Y = true; };
template <> struct ModeT<MODE_ABS_J> { static constexpr int nspec = 2; static constexpr bool J = true, Y = false; };
template <> struct ModeT<MODE_ABS_H> { static constexpr int nspec = 2; static constexpr bool J = true, Y = true; };

// Floats of coefficient record per station and frequency.  itype 0: A'_s per
// spectrum.  itype 1, per spectrum: the hat record (ah, ph, th, tt) in the
// station's record, b_s in a second plane after the nrp records and the
// kappa record in a third (see station_coef and coef_store).
FJ_HD constexpr int coef_count(int mode, int itype) {
  return ((mode == MODE_J || mode == MODE_Y) ? 1 : 2) * (itype == 0 ? 1 : 6);
}
// Floats of the coefficients of one frequency: nrp records (for itype 1
// with the b and kappa planes) and a header of 8 floats at coef_header:
// [0] the long-segment limit (long_segment_limit; itype 1), [1] the anchor
// limit (anchor_limit).
FJ_HD constexpr size_t coef_header(int nrp, int mode, int itype) {
  return (size_t)nrp * coef_count(mode, itype);
}
FJ_HD constexpr size_t coef_block(int nrp, int mode, int itype) {
  return coef_header(nrp, mode, itype) + 8;
}
// itype 1 offsets inside a frequency's coefficients (ns spectra): the b
// plane [nrp][ns], the kappa plane [ns][nrp], the header.
FJ_HD constexpr size_t lin_b_offset(int nrp, int ns) { return (size_t)nrp * 4 * ns; }
FJ_HD constexpr size_t lin_kappa_offset(int nrp, int ns) { return (size_t)nrp * 5 * ns; }
FJ_HD constexpr size_t lin_header(int nrp, int ns) { return (size_t)nrp * 6 * ns; }

// itype 1: segments with k_max h > kLongDx at some frequency ("long": k_max
// = 2 pi |f| / min |c| over the velocity grid) are left out of the (a, p, t)
// records of that frequency and summed per lane instead.  For the others
// the rule's terms in the records stay within about 1.3 times g even where
// a lane replaces them by the G difference, so that replacement cancels
// nothing large.  (A limit of kTau, which sends every segment on which a
// lane may take the G difference to the per-lane sum, was slower on small
// arrays too: 4096 ragged sub-arrays 672 ms against 645 ms.)
constexpr float kLongDx = 2.0f;

// h_limit of a frequency: a segment is long when its float gap h has
// f32_bits(h) > f32_bits(h_limit) (gaps are non-negative; +inf at f = 0).
FJ_HD inline float long_segment_limit(float f, float cmin) {
  const double k = kTwoPi * fabs((double)f) / (double)cmin;
  return k > 0.0 ? (float)((double)kLongDx / k) : INFINITY;
}

// Anchored phases (block_phases): a block of 4 stations whose offsets from
// its anchor satisfy |q dr| <= kAnchorRev for every velocity of the grid,
// i.e. max |dr| <= kAnchorRev min|c| / |f|, the anchor limit of frequency
// f (+inf at f = 0).  It depends on the frequency and the grid only, so
// every kernel that evaluates a term takes the same phase for it.
constexpr float kAnchorRev = 1.0f;
// Anchored phases are taken in the one-spectrum modes.  Measured on the
// MI355X at 19,900 pairs x 4096 f x 3000 c: trapezoid J 82.0 -> 76.4 ms,
// linear-segment J 177.3 -> 173.4 ms; the two-spectrum kernels hold more
// station data per block, and the offsets spilled their scalar registers
// (trapezoid J - Y 119 -> 183 ms, linear-segment J - Y 305 -> 321 ms).
template <int MODE>
struct Anchored {
  static constexpr bool on = (MODE == MODE_J || MODE == MODE_Y);
};
FJ_HD inline float anchor_limit(float f, float cmin) {
  const double af = fabs((double)f);
  return af > 0.0 ? (float)((double)kAnchorRev * (double)cmin / af) : INFINITY;
}

// Occupancy floor of the F-J kernels (waves per SIMD; the compiler then
// keeps VGPRs within 512 / waves).  The dense linear-segment kernels are
// compiled without machine loop-invariant code motion
// (fj_kernels_linseg.hip) and need no floor there: J - Y takes 62 VGPRs
// and 101 SGPRs (7 waves), J 43 VGPRs (8 waves), with no scratch.  With
// that code motion on, J - Y took 217 VGPRs (2 waves) unheld and ran 16%
// faster held to 3 (168 VGPRs, 116 bytes of spills per lane): 19,900 pairs
// x 4096 f x 3000 c in 392 ms, against 323 ms now.  Measured without the
// code motion, occupancy limited by LDS: J - Y 451 / 371 / 353 / 332 ms at
// 3 / 4 / 5 / 6 waves, 323 ms at its natural 7; J 257 / 241 ms at 5 / 6,
// 222 ms at 8.  The persistent batched linear-segment J kernel is held to
// 6 waves (4096 ragged sub-arrays in 604 ms against 631 ms at 5 and 674 ms
// at 4); the item-list kernel (ITEMS) holds 4: the learned bands of a
// 300-pair field in 9.0 ms against 9.5 ms at 5 and 11.6 ms at 6.
// -DFJ_OCC_I1_JMY=N / -DFJ_OCC_I1_J=N / -DFJ_OCC_I1_J_ITEMS=N override for
// tuning runs.
#ifndef FJ_OCC_I1_JMY
#define FJ_OCC_I1_JMY 3
#endif
#ifndef FJ_OCC_I1_J
#define FJ_OCC_I1_J 6
#endif
#ifndef FJ_OCC_I1_J_ITEMS
#define FJ_OCC_I1_J_ITEMS 4
#endif
// How the linear-segment kernels hold the small-argument (x < 8) double
// polynomials.  Measured on the MI355X (J kernel, 19,900 pairs x 512 f x
// 3000 c): literals 39.6 ms; a __constant__ table 73 ms, because the
// compiler hoists its loads into scalar registers that spill to vector
// lanes and are read back in the station loop; the table re-read inside
// the small-argument branch (address hidden from hoisting) 58 ms, because
// the station stream through the scalar cache evicts it between uses.  The
// trapezoid kernels have the registers to spare and take the table (5-13%
// faster than literals).  -DFJ_FIT_I1=1 selects the table, for tuning runs.
#ifndef FJ_FIT_I1
#define FJ_FIT_I1 FIT_LITERAL
#endif
template <int MODE, int ITYPE, bool ITEMS = false>
struct FjOcc {
  static constexpr int waves =
      ITYPE == 0 ? 1
      : MODE == MODE_J_MINUS_Y ? FJ_OCC_I1_JMY
      : MODE == MODE_J ? (ITEMS ? FJ_OCC_I1_J_ITEMS : FJ_OCC_I1_J)
      : 1;
};

// ---- coefficients of station s (one frequency) ----------------------------
// g_prev/g_cur/g_next: spectrum values at stations s-1, s, s+1 (only those
// that exist are read).  rp, rs, rn: r_{s-1}, r_s, r_{s+1} (likewise).
// itype 0 writes A'_s per spectrum.  itype 1 writes, per spectrum m, the
// hat record rec[m] = a_s sqrt r_s, rec[NSPEC + m] = p_s sqrt r_s,
// rec[2 NSPEC + m] = t_s / sqrt r_s, rec[3 NSPEC + m] = t_s sqrt r_s (each
// formed in double and rounded once; see the hat form before far_units),
// rec[4 NSPEC + m] = b_s and rec[5 NSPEC + m] = kappa_of(rec[m],
// station_isr(r_s)).  a_s, p_s and t_s are
//   the station's share of sum_s alpha_s g_s - sum_s b_s E_s with every E_s
//   (= 2 D_s, segment s between stations s-1 and s) taken by the
//   Euler-Maclaurin rule,
//     E_s = g_{s-1} + g_s - (k h_s / 6)(g'_s - g'_{s-1})
//           + (k^3 h_s^3 / 360)(g'''_s - g'''_{s-1}),
//   collected per station:  a_s g_s + k p_s g'_s + k^3 t_s g'''_s with
//     a_s = alpha_s - b_s - b_{s+1},
//     p_s = (b_s h_s - b_{s+1} h_{s+1}) / 6,
//     t_s = -(b_s h_s^3 - b_{s+1} h_{s+1}^3) / 360,
//   b_s = beta_s / 2 (zero for s = 0 and past the last station), h_s the
//   float gap r_s - r_{s-1} as in the station arrays.  A lane whose segment
//   s takes the G difference instead adds b_s (E_s^rule - E_s^G).  Long
//   segments (h_s beyond h_limit, long_segment_limit) are left out of a, p
//   and t; the lanes add their -b_s E_s themselves.
template <int ITYPE, int NSPEC>
FJ_HD inline void station_coef(int s, int nr, float rp, float rs, float rn,
                               const float* g_prev, const float* g_cur, const float* g_next,
                               float h_limit, float* rec) {
  if (ITYPE == 0) {
    // trapezoid weight (src/FJcpu.cpp:354-356): each segment gives d/2 to
    // both of its end stations, d = max(h, 0.1) in float as the reference.
    // A_s = W_s r_s U_s, stored divided by sqrt(r_s) (see trapz_lane).
    double w = 0.0;
    if (s > 0) w += (double)fmaxf(rs - rp, 0.1f);
    if (s + 1 < nr) w += (double)fmaxf(rn - rs, 0.1f);
    const double wr = 0.5 * w * sqrt((double)rs);
    for (int k = 0; k < NSPEC; k++) rec[k] = (float)(wr * (double)g_cur[k]);
  } else {
    // slope over the clamped gap (src/FJcpu.cpp:259-261): b = (g2-g1)/dr0,
    // dr0 = max(h, 0.1); the segment's line reaches g1 + b h at its right end.
    const bool left = s > 0, right = s + 1 < nr;
    const uint32_t hb = f32_bits(h_limit);
    const bool inl = left && f32_bits(rs - rp) <= hb;    // segment s in the records
    const bool inr = right && f32_bits(rn - rs) <= hb;   // segment s + 1 in the records
    const double hl = left ? (double)(rs - rp) : 0.0;
    const double hr = right ? (double)(rn - rs) : 0.0;
    const double ratl = left ? hl / (double)fmaxf(rs - rp, 0.1f) : 0.0;
    const double ratr = right ? hr / (double)fmaxf(rn - rs, 0.1f) : 0.0;
    const double sq = sqrt((double)rs), isq = 1.0 / sq;
    for (int k = 0; k < NSPEC; k++) {
      const double g2 = g_cur[k];
      const double g1 = left ? (double)g_prev[k] : 0.0;
      const double g3 = right ? (double)g_next[k] : 0.0;
      const double bl = left ? 0.5 * (g2 - g1) * ratl : 0.0;   // b_s
      const double br = right ? 0.5 * (g3 - g2) * ratr : 0.0;  // b_{s+1}
      const double el = left ? g1 + 2.0 * bl : 0.0;  // left limit at r_s
      const double er = right ? g2 : 0.0;             // right limit at r_s
      const double alpha = el - er;
      const double cl = inl ? bl : 0.0, cr = inr ? br : 0.0;
      const double a = alpha - cl - cr;
      const double p = (cl * hl - cr * hr) / 6.0;
      const double t = -(cl * hl * hl * hl - cr * hr * hr * hr) / 360.0;
      rec[k] = (float)(a * sq);
      rec[NSPEC + k] = (float)(p * sq);
      rec[2 * NSPEC + k] = (float)(t * isq);
      rec[3 * NSPEC + k] = (float)(t * sq);
      rec[4 * NSPEC + k] = (float)bl;
      rec[5 * NSPEC + k] = kappa_of(rec[k], station_isr(rs));
    }
  }
}

// itype 0, J - Y: the station's two coefficients (c0 for the J sum of U,
// c1 for the Y sum of its Hilbert transform) in polar form, R = |c0 + i c1|
// and psi = arg(c0 + i c1) in revolutions.  Then c0 J0 - c1 Y0 = amp R M0
// cos(th + phi0 + psi) in the far form: one cosine a term instead of a sine
// and a cosine (trapz_far); the other argument ranges rotate the phase by
// psi as well, or rebuild c0 and c1 below x = 8.
FJ_HD constexpr bool polar_records(int mode, int itype) {
  return itype == 0 && mode == MODE_J_MINUS_Y;
}
FJ_HD inline void to_polar(float* rec) {
  const double c0 = rec[0], c1 = rec[1];
  rec[0] = (float)sqrt(c0 * c0 + c1 * c1);
  rec[1] = (float)(atan2(c1, c0) / kTwoPi);
}

// Stores station s's record into the coefficients of one frequency (cf,
// nrp stations): itype 0 at cf[s NSPEC]; itype 1 the hat record at
// cf[s 4 NSPEC], b_s at cf[lin_b_offset + s NSPEC + m] and kappa at
// cf[lin_kappa_offset + m nrp + s].
template <int ITYPE, int NSPEC>
FJ_HD inline void coef_store(float* cf, int nrp, int s, const float* rec) {
  if (ITYPE == 0) {
    for (int k = 0; k < NSPEC; k++) cf[(size_t)s * NSPEC + k] = rec[k];
  } else {
    for (int k = 0; k < 4 * NSPEC; k++) cf[(size_t)s * 4 * NSPEC + k] = rec[k];
    for (int k = 0; k < NSPEC; k++) {
      cf[lin_b_offset(nrp, NSPEC) + (size_t)s * NSPEC + k] = rec[4 * NSPEC + k];
      cf[lin_kappa_offset(nrp, NSPEC) + (size_t)k * nrp + s] = rec[5 * NSPEC + k];
    }
  }
}

// ---- itype 0 ---------------------------------------------------------------
// Coefficients A'_s = A_s / sqrt(r_s) (NC = nspec floats per station); the
// station adds A'_s Z0(x_s) / amp_s with amp_s = sqrt(2/(pi x_s)) the
// Hankel amplitude, i.e. P0 cos - Q0 sin for x >= 8, and the lane
// multiplies the sum by sqrt(2/(pi k)) = amp_s sqrt(r_s).  At k = 0 (f = 0)
// there is no amplitude: the station adds A'_s Z0 sqrt(r_s), the lane
// multiplies by 1.
// Station term A'_s (vj, vy) added to the lane's sums.
template <int MODE>
FJ_HD inline void trapz_add(const float* c, float vj, float vy, float& pR, float& pI) {
  if (MODE == MODE_J) {
    pR = __builtin_fmaf(c[0], vj, pR);
  } else if (MODE == MODE_Y) {
    pR = __builtin_fmaf(c[0], vy, pR);
  } else if (MODE == MODE_J_MINUS_Y) {
    pR = __builtin_fmaf(c[0], vj, pR);
    pR = __builtin_fmaf(-c[1], vy, pR);
  } else if (MODE == MODE_ABS_J) {
    pR = __builtin_fmaf(c[0], vj, pR);
    pI = __builtin_fmaf(c[1], vj, pI);
  } else {  // MODE_ABS_H: (FJ[re]-FY[im]) + i (FJ[im]+FY[re])
    pR = __builtin_fmaf(c[0], vj, pR);
    pR = __builtin_fmaf(-c[1], vy, pR);
    pI = __builtin_fmaf(c[1], vj, pI);
    pI = __builtin_fmaf(c[0], vy, pI);
  }
}

// x >= kLaneFar, z = 8/x, ph = the station's reduced phase (block_phases):
// J0 / amp and Y0 / amp in the modulus-phase form J0 = amp M0 cos(th + phi),
// Y0 = amp M0 sin(th + phi), one transcendental per basis.  Far blocks and
// the per-lane path both use it, so a term does not depend on which lanes
// share its wave.
template <int MODE>
FJ_HD inline void far_unit(float ph, float z, float& vj, float& vy) {
  using M = ModeT<MODE>;
  const float t = z * z;
  const float m0 = fitm_M0(t);
  const float u = __builtin_fmaf(z, fitm_ph0(t), ph);
  if (M::J && M::Y) {
    float s, co;
    sincos_rev(u, &s, &co);
    vj = m0 * co;
    vy = m0 * s;
  } else if (M::J) {
    vj = m0 * cos_rev(u);
  } else {
    vy = m0 * sin_rev(u);
  }
}

// J - Y with polar records (c = R, psi): R M0 cos(u + psi), x >= kLaneFar.
FJ_HD inline float far_polar(float ph, float z, float psi) {
  const float t = z * z;
  const float u = __builtin_fmaf(z, fitm_ph0(t), ph + psi);
  return fitm_M0(t) * cos_rev(u);
}

// One station of the trapezoid sum: rs = r_s, ph = its reduced phase
// (block_phases; OWN_PH: formed here, the same value, for far lanes only),
// ir = 1/r_s, c = its record.
template <int MODE, bool OWN_PH = false>
FJ_HD inline void trapz_station(double rs, float ph, float ir, const LaneK& L, const float* c,
                                float& pR, float& pI) {
  using M = ModeT<MODE>;
  float vj = 0.0f, vy = 0.0f;
  const float z = L.zk * ir;  // 8 / x; inf when k = 0 or r = 0
  // x >= kLaneFar: the lane's own far test, implied by the block's
  const bool far = f32_bits(ir) <= L.irf;
  if (OWN_PH && far) ph = phase_rev(L.q, rs);
  if (MODE == MODE_J_MINUS_Y && far) {   // polar records
    pR = __builtin_fmaf(c[0], far_polar(ph, z, c[1]), pR);
    return;
  }
  if (far) {
    far_unit<MODE>(ph, z, vj, vy);
  } else {
    const double xd = L.q2pi * rs;
    double y, Lg;
    const Bess b = bessel_small<false, M::Y, FIT_TABLE>(xd, &y, &Lg);
    const float ia = L.kf > 0.0f ? sqrt_f(kPiOver2 * (float)xd) : sqrt_f((float)rs);
    vj = b.j0 * ia;
    vy = b.y0 * ia;
  }
  if (MODE == MODE_J_MINUS_Y) {   // polar records: c0 = R cos psi, c1 = R sin psi
    float s, co;
    sincos_rev(c[1], &s, &co);
    pR = __builtin_fmaf(c[0] * co, vj, pR);
    pR = __builtin_fmaf(-(c[0] * s), vy, pR);
    return;
  }
  trapz_add<MODE>(c, vj, vy, pR, pI);
}

// x >= kLaneFar on every lane (a far block): no branch on the argument.
template <int MODE>
FJ_HD inline void trapz_far(float ph, float ir, const LaneK& L, const float* c, float& pR,
                            float& pI) {
  if (MODE == MODE_J_MINUS_Y) {   // polar records
    pR = __builtin_fmaf(c[0], far_polar(ph, L.zk * ir, c[1]), pR);
    return;
  }
  float vj = 0.0f, vy = 0.0f;
  far_unit<MODE>(ph, L.zk * ir, vj, vy);
  trapz_add<MODE>(c, vj, vy, pR, pI);
}

// Reduced phases u_s = frac(q r_s - 1/8) (revolutions) of the NB stations
// from sb.  Anchored blocks (ANCH: max |dr| <= the frequency's anchor
// limit) reduce in double once per block of 4, at its first station a, and
// add q dr_s in float: u_s = u_a + q_f dr_s with |q dr_s| <= kAnchorRev, so
// that u_s is off by at most 2^-25 (u_a) + 2^-24 kAnchorRev (q_f) +
// 2^-24 (the sum, below 2) = 1.5e-7 revolutions.  Other blocks reduce each
// station in double (error 2^-25).  Either way one phase per station and
// block, used by every kind of block and lane path alike.
template <int NB>
FJ_HD inline void direct_phases(const double* rv, const LaneK& L, float* u) {
  double w[NB];
  // in three stages of NB independent instructions (latency)
#pragma unroll
  for (int k = 0; k < NB; k++) w[k] = __builtin_fma(L.q, rv[k], -0.125);
#pragma unroll
  for (int k = 0; k < NB; k++) w[k] = fract_d(w[k]);
#pragma unroll
  for (int k = 0; k < NB; k++) u[k] = (float)w[k];
}
template <int NB>
FJ_HD inline void block_phases(const Stations& S, int sb, const LaneK& L, bool anch, float* u) {
  if (anch) {
    float ua[NB / 4];
#pragma unroll
    for (int a = 0; a < NB / 4; a++) ua[a] = phase_rev(L.q, S.rr[sb + 4 * a]);
#pragma unroll
    for (int k = 0; k < NB; k++) u[k] = __builtin_fmaf(L.qf, S.dr[sb + k], ua[k >> 2]);
  } else {
    double rv[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) rv[k] = S.rr[sb + k];
    direct_phases<NB>(rv, L, u);
  }
}

// A far block of kUnroll stations evaluated stage by stage (all fits, then
// all cosines, then the sums in station order), so that each stage offers
// kUnroll independent instructions to hide the latency of the
// transcendental.  The same operations as trapz_far per station, summed in
// the same order: bit-identical results.  Measured on the MI355X at 19,900
// pairs x 4096 f x 3000 c: J 104.6 ms against 107.3 ms station by station,
// J - Y 153.6 against 157.3 ms.
template <int MODE, int NB>
FJ_HD inline void trapz_far_block(const float* ph, const float* iv, const LaneK& L,
                                  const float* cv, float& pR, float& pI) {
  constexpr int NC = ModeT<MODE>::nspec;
  constexpr int kUnroll = NB;
  if (MODE == MODE_J || MODE == MODE_J_MINUS_Y) {
    float z[kUnroll], t[kUnroll], u[kUnroll], v[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; k++) {
      z[k] = L.zk * iv[k];
      t[k] = z[k] * z[k];
    }
#pragma unroll
    for (int k = 0; k < kUnroll; k++) {
      const float p = MODE == MODE_J_MINUS_Y ? ph[k] + cv[k * NC + 1] : ph[k];
      u[k] = __builtin_fmaf(z[k], fitm_ph0(t[k]), p);
    }
#pragma unroll
    for (int k = 0; k < kUnroll; k++) v[k] = fitm_M0(t[k]) * cos_rev(u[k]);
#pragma unroll
    for (int k = 0; k < kUnroll; k++) pR = __builtin_fmaf(cv[k * NC], v[k], pR);
  } else {
#pragma unroll
    for (int k = 0; k < kUnroll; k++) trapz_far<MODE>(ph[k], iv[k], L, cv + k * NC, pR, pI);
  }
}

// Stations per block of the trapezoid loop, per mode.
template <int MODE>
struct TrapzBlock {
  static constexpr int n = Anchored<MODE>::on ? kUnroll : 4;
};

template <int MODE>
FJ_HD inline float trapz_lane(const Stations& S, const float* __restrict__ cf, int nrp, double q) {
  constexpr int NC = ModeT<MODE>::nspec;
  constexpr int TB = TrapzBlock<MODE>::n;
  const LaneK L = lane_consts(q);
  const uint32_t irfar = far_threshold(L.q2pi);
  const uint32_t alim = f32_bits(cf[coef_header(nrp, MODE, 0) + 1]);
  double accR = 0.0, accI = 0.0;
  for (int s0 = 0; s0 < nrp; s0 += kPartial) {
    const int s1 = nrp < s0 + kPartial ? nrp : s0 + kPartial;
    float pR = 0.0f, pI = 0.0f;
    for (int sb = s0; sb < s1; sb += TB) {
      float iv[TB];
      float cv[TB * NC];
#pragma unroll
      for (int k = 0; k < TB; k++) iv[k] = S.ir[sb + k];
#pragma unroll
      for (int k = 0; k < TB * NC; k++) cv[k] = cf[sb * NC + k];
      // every station of the block has x >= kLaneFar (block words)
      const bool far = S.bw[TB == 8 ? sb + 2 : sb] <= irfar;
      if (Anchored<MODE>::on) {
        float ph[TB];
        block_phases<TB>(S, sb, L, S.bw[TB == 8 ? sb + 6 : sb + 3] <= alim, ph);
        if (far) {
          trapz_far_block<MODE, TB>(ph, iv, L, cv, pR, pI);
        } else {
#pragma unroll
          for (int k = 0; k < TB; k++)
            trapz_station<MODE>(S.rr[sb + k], ph[k], iv[k], L, cv + k * NC, pR, pI);
        }
      } else {
        // the other modes form the phases in the far path only (in a block
        // that is not far, for the lanes that are), from the distances
        // loaded with the block: fewer scalar registers held across the
        // branch (J - Y 123 -> 119 ms)
        double rv[TB];
#pragma unroll
        for (int k = 0; k < TB; k++) rv[k] = S.rr[sb + k];
        if (far) {
          float ph[TB];
          direct_phases<TB>(rv, L, ph);
          trapz_far_block<MODE, TB>(ph, iv, L, cv, pR, pI);
        } else {
#pragma unroll
          for (int k = 0; k < TB; k++)
            trapz_station<MODE, true>(rv[k], 0.0f, iv[k], L, cv + k * NC, pR, pI);
        }
      }
    }
    accR += (double)pR;
    accI += (double)pI;
  }
  const double fin = L.kf > 0.0f ? sqrt(2.0 / (kPiD * L.q2pi)) : 1.0;
  if (MODE == MODE_ABS_J || MODE == MODE_ABS_H) return (float)(sqrt(accR * accR + accI * accI) * fin);
  return (float)(accR * fin);
}

// ---- itype 1 ---------------------------------------------------------------
// Y basis: x Y1(x) -> -2/pi as x -> 0.  Because sum_s alpha_s = sum_s beta_s
// exactly (the jumps and rises of one interpolant telescope), that constant
// drops out of alpha g - beta D; the Y terms therefore use the shifted
// g~ = x Y1 + 2/pi and D~ = D + 2/pi, which removes a cancellation of up to
// ~1/(x^2 ln x) at small x.  G is stored shifted (G + 2x/pi) where it comes
// from the small-argument branch (sh = 1) and unshifted otherwise.
// E = 2 D (and 2 D~) with b = beta / 2 (station_coef); the rule's form of
// every E is collected per station into the records, so the station loop
// adds a few products per spectrum and evaluates E itself only on the
// segments that take the G difference.
struct Ends {  // per-basis values at a station
  float g, gp, g3, G, sh;
};

// 2/dx int_{x_a}^{x_b} g by Euler-Maclaurin with two end corrections;
// e1 = -dx/6, e3 = dx^3/360 are shared by both bases.
FJ_HD inline float em_divided2(const Ends& a, const Ends& b, float e1, float e3) {
  float d = a.g + b.g;
  d = __builtin_fmaf(e1, b.gp - a.gp, d);
  d = __builtin_fmaf(e3, b.g3 - a.g3, d);
  return d;
}

// 2 D~ for the Y basis from stored (possibly shifted) G values; inv2 = 2/dx.
FJ_HD inline float gdiff_y2(const Ends& a, float xa, const Ends& b, float xb, float inv2) {
  if (a.sh == b.sh)
    return __builtin_fmaf(b.G - a.G, inv2, a.sh != 0.0f ? 0.0f : 2.0f * kTwoOverPi);
  const float Ga = a.sh != 0.0f ? a.G : __builtin_fmaf(kTwoOverPi, xa, a.G);
  const float Gb = b.sh != 0.0f ? b.G : __builtin_fmaf(kTwoOverPi, xb, b.G);
  return (Gb - Ga) * inv2;
}

// ---- far stations: the hat form --------------------------------------------
// For x >= kLaneFar the Bessel functions are written over their Hankel
// amplitude amp = sqrt(2/(pi x)) = ak / sqrt(r), ak = sqrt(2/(pi k)):
//   U0 = J0 / amp, U1 = J1 / amp, V0 = Y0 / amp, V1 = Y1 / amp.
// With g = x Z1, g' = x Z0, g''' = -(Z1 + x Z0) and x = k r, a station's
// record terms become
//   a g + k p g' + k^3 t g''' = k ak [(ah - k^2 th) U1 + k (ph - k^2 tt) U0]
// with the hat records ah = a sqrt r, ph = p sqrt r, th = t / sqrt r and
// tt = t sqrt r (station_coef).  The lane keeps four float sums, of
// ah U1, th U1, ph U0 and tt U0 (one product each: a multiply-add reads
// one scalar operand), scaled by k ak = sqrt(2k/pi), -k^2 times that, k
// times that and -k^3 times that when flushed: no x, no amplitude and no g
// products per station.  The Y basis's shifted
// g~ = x Y1 + 2/pi adds (2/pi) a_s on top, the kappa record (kappa_of).
// Every far station takes this form, in every kind of block, so a
// station's terms do not depend on which lanes share its wave.
struct FarUnits {
  float u0, u1, v0, v1;
};

// Bases with J only: the modulus-phase form, one transcendental per order
// (U0 = M0 cos(th + phi0), U1 = M1 sin(th + phi1), th = x - pi/4).  With
// Y: one sine/cosine pair of th through the P, Q fits serves all four.
template <bool NY>
FJ_HD inline FarUnits far_units(float u, float z) {
  const float t = z * z;
  FarUnits w;
  if (!NY) {
    w.u0 = fitm_M0(t) * cos_rev(__builtin_fmaf(z, fitm_ph0(t), u));
    w.u1 = fitm_M1(t) * sin_rev(__builtin_fmaf(z, fitm_ph1(t), u));
    w.v0 = w.v1 = 0.0f;
  } else {
    float sn, co;
    sincos_rev(u, &sn, &co);
    const float P0 = fitm_P0(t), Q0 = z * fitm_q0(t);
    const float P1 = fitm_P1(t), Q1 = z * fitm_q1(t);
    w.u0 = __builtin_fmaf(P0, co, -Q0 * sn);
    w.v0 = __builtin_fmaf(P0, sn, Q0 * co);
    w.u1 = __builtin_fmaf(P1, sn, Q1 * co);
    w.v1 = __builtin_fmaf(Q1, sn, -P1 * co);
  }
  return w;
}

template <int MODE>
struct LinState {
  Ends pj = {0.f, 0.f, 0.f, 0.f, 0.f}, py = {0.f, 0.f, 0.f, 0.f, 0.f};
  float xprev = 0.0f;
  bool gnext = false;  // the segment right of the previous station uses the G difference
  bool lnext = false;  // ... is long (not in the records; uniform)
  // float partial sums of the near stations' a g, p g' and t g''' terms and
  // of the corrections (R and, for the modulus modes, I) ...
  float aR = 0.0f, aI = 0.0f, pR = 0.0f, pI = 0.0f, tR = 0.0f, tI = 0.0f;
  // ... and of the far stations' hat form: ah U1, th U1, ph U0, tt U0
  float haR = 0.0f, haI = 0.0f, htR = 0.0f, htI = 0.0f, hpR = 0.0f, hpI = 0.0f, hqR = 0.0f,
        hqI = 0.0f;
};

// R (+= w . v) and I for one quantity v per basis (vJ, vY) and one
// coefficient per spectrum (w[0], w[1]):
//   J: R = w0 vJ.   Y: R = w0 vY.   J - Y: R = w0 vJ - w1 vY.
//   |J| (Re, Im spectra): R = w0 vJ, I = w1 vJ.
//   |H| = |sum U (J + iY)|: R = w0 vJ - w1 vY, I = w1 vJ + w0 vY.
template <int MODE>
FJ_HD inline void lin_acc(const float* w, float vJ, float vY, float& R, float& I) {
  if (MODE == MODE_J) {
    R = __builtin_fmaf(w[0], vJ, R);
  } else if (MODE == MODE_Y) {
    R = __builtin_fmaf(w[0], vY, R);
  } else if (MODE == MODE_J_MINUS_Y) {
    R = __builtin_fmaf(w[0], vJ, R);
    R = __builtin_fmaf(-w[1], vY, R);
  } else if (MODE == MODE_ABS_J) {
    R = __builtin_fmaf(w[0], vJ, R);
    I = __builtin_fmaf(w[1], vJ, I);
  } else {  // MODE_ABS_H
    R = __builtin_fmaf(w[0], vJ, R);
    R = __builtin_fmaf(-w[1], vY, R);
    I = __builtin_fmaf(w[1], vJ, I);
    I = __builtin_fmaf(w[0], vY, I);
  }
}

// The (2/pi) a_s of the shifted Y terms, kap[m] per spectrum, with the
// signs lin_acc gives vY.
template <int MODE>
FJ_HD inline void kappa_acc(const float* kap, float& R, float& I) {
  if (MODE == MODE_Y) {
    R += kap[0];
  } else if (MODE == MODE_J_MINUS_Y) {
    R -= kap[1];
  } else if (MODE == MODE_ABS_H) {
    R -= kap[1];
    I += kap[0];
  }
}

// A far station's record terms (hat records h: ah, ph, th, tt, nspec floats
// each) into the four hat sums.
template <int MODE>
FJ_HD inline void far_acc(const float* h, const FarUnits& w, LinState<MODE>& st) {
  constexpr int NS = ModeT<MODE>::nspec;
  lin_acc<MODE>(h, w.u1, w.v1, st.haR, st.haI);
  lin_acc<MODE>(h + 2 * NS, w.u1, w.v1, st.htR, st.htI);
  lin_acc<MODE>(h + NS, w.u0, w.v0, st.hpR, st.hpI);
  lin_acc<MODE>(h + 3 * NS, w.u0, w.v0, st.hqR, st.hqI);
}

// One station of the linear-segment sum.  ph = its reduced phase
// (block_phases), dxL = k (r_s - r_{s-1}),
// dxR = k (r_{s+1} - r_s) (0 past the last station), hR = r_{s+1} - r_s,
// hlim = the bits of the frequency's long-segment limit, h = the station's
// hat record, kap = its kappa record (smooth blocks; the others form it),
// cb = its b record.  FAR: x >= kLaneFar on every lane of the wave.  SMOOTH
// (with FAR): no segment at this station is long or takes the G
// difference on any lane, so no per-station test is made.
template <int MODE, bool FAR, bool SMOOTH = false>
FJ_HD inline void linseg_station(int s, double rs, float ph, float ir, float isr, float dxL, float dxR,
                                 float hR, uint32_t hlim, const LaneK& L, const float* h,
                                 const float* kap, const float* cb,
                                 LinState<MODE>& st) {
  using M = ModeT<MODE>;
  constexpr bool NJ = M::J, NY = M::Y;
  constexpr int NS = M::nspec;
  const float z = L.zk * ir;   // 8 / x; inf when k = 0 or r = 0
  if (SMOOTH) {
    far_acc<MODE>(h, far_units<NY>(ph, z), st);
    if (NY) kappa_acc<MODE>(kap, st.aR, st.aI);
    return;
  }
  const double xd = L.q2pi * rs;
  const float x = (float)xd;
  const bool useG_L = s > 0 && st.gnext;
  const bool longL = st.lnext;
  st.lnext = f32_bits(hR) > hlim;
  // in a far block both ends of every segment have x >= kLaneFar, where the
  // rule's threshold min(kTau, kKappa min(x_a, x_b)) is kTau
  const bool gR = FAR ? fabsf(dxR) >= kTau : use_gdiff_right(x, dxR);
  st.gnext = gR;
  const bool needG = useG_L || gR;
  Ends cj = {0.f, 0.f, 0.f, 0.f, 0.f}, cy = {0.f, 0.f, 0.f, 0.f, 0.f};
  // the lane's own far test, implied by the block's (far_threshold)
  if (FAR || f32_bits(ir) <= L.irf) {
    const FarUnits w = far_units<NY>(ph, z);
    far_acc<MODE>(h, w, st);
    if (NY) {
      float kp[NS];
#pragma unroll
      for (int m = 0; m < NS; m++) kp[m] = kappa_of(h[m], isr);
      kappa_acc<MODE>(kp, st.aR, st.aI);
    }
    // the ends, for a correction of this segment or the next
    if (needG || longL || st.lnext) {
      const float amp = L.ak * isr;
      const float j0 = amp * w.u0, j1 = amp * w.u1, y0 = amp * w.v0, y1 = amp * w.v1;
      if (NJ) {
        cj.g = x * j1;
        cj.gp = x * j0;
        cj.g3 = -(j1 + cj.gp);
      }
      if (NY) {
        cy.g = __builtin_fmaf(x, y1, kTwoOverPi);
        cy.gp = x * y0;
        cy.g3 = -(y1 + cy.gp);
      }
      if (needG) {
        const float t = z * z;
        const float r0 = fit_R0(t), xr1 = x * fit_R1(t);
        if (NJ) cj.G = __builtin_fmaf(r0, j1, __builtin_fmaf(-xr1, j0, 1.0f));
        if (NY) cy.G = __builtin_fmaf(r0, y1, -xr1 * y0);
      }
    }
  } else {
    // x < kLaneFar: the double-precision small-argument polynomials
    double y, Lg;
    const Bess b = bessel_small<true, NY, FJ_FIT_I1>(xd, &y, &Lg);
    const float j0 = b.j0, j1 = b.j1, y0 = b.y0, y1 = b.y1, gy = b.gy;
    if (needG) g_small<FJ_FIT_I1>(xd, y, Lg, NJ, NY, &cj.G, &cy.G);
    cy.sh = 1.0f;
    if (NJ) {
      cj.g = x * j1;
      cj.gp = x * j0;
      cj.g3 = -(j1 + cj.gp);
    }
    if (NY) {
      cy.g = gy;
      cy.gp = x * y0;
      cy.g3 = -(y1 + cy.gp);
    }
    // the (a, p, t) records back from the hat records: a = ah / sqrt r,
    // p = ph / sqrt r, t = tt / sqrt r (isr is finite: station_isr)
    float ca[NS], cp[NS], ct[NS];
#pragma unroll
    for (int m = 0; m < NS; m++) {
      ca[m] = h[m] * isr;
      cp[m] = h[NS + m] * isr;
      ct[m] = h[3 * NS + m] * isr;
    }
    lin_acc<MODE>(ca, cj.g, cy.g, st.aR, st.aI);
    lin_acc<MODE>(cp, cj.gp, cy.gp, st.pR, st.pI);
    lin_acc<MODE>(ct, cj.g3, cy.g3, st.tR, st.tI);
  }
  // Two rare cases need the segment to the left (a wave-uniform branch; the
  // b record cb is used only there): a lane whose segment takes the G
  // difference adds b (E^rule - E^G) per basis, and a long segment (longL,
  // uniform) adds -b E on every lane, E by the rule or the G difference as
  // the lane chose.
  if (s > 0 && (longL || wave_any(useG_L))) {
    float dJ = 0.0f, dY = 0.0f;
    if (useG_L) {   // per lane: the G difference (and the rule, for a short segment)
      const float inv2 = rcp_f(0.5f * dxL);
      if (NJ) dJ = -((cj.G - st.pj.G) * inv2);
      if (NY) dY = -gdiff_y2(st.py, st.xprev, cy, x, inv2);
      if (!longL) {   // uniform
        const float e1 = dxL * (-1.0f / 6.0f);
        const float e3 = dxL * dxL * dxL * (1.0f / 360.0f);
        if (NJ) dJ += em_divided2(st.pj, cj, e1, e3);
        if (NY) dY += em_divided2(st.py, cy, e1, e3);
      }
    } else if (longL) {   // the rule, on the lanes of a long segment without G
      const float e1 = dxL * (-1.0f / 6.0f);
      const float e3 = dxL * dxL * dxL * (1.0f / 360.0f);
      if (NJ) dJ = -em_divided2(st.pj, cj, e1, e3);
      if (NY) dY = -em_divided2(st.py, cy, e1, e3);
    }
    lin_acc<MODE>(cb, dJ, dY, st.aR, st.aI);
  }
  st.xprev = x;
  st.pj = cj;
  st.py = cy;
}

// Far blocks in the linear-segment loop that still test every station (a
// long segment or a G difference somewhere in the block).  Measured on the
// MI355X they pay for the two-spectrum modes (J - Y: 69 -> 58 ms per 512
// dense columns) and not for the one-spectrum modes, whose general loop is
// lighter in registers; -DFJ_FAR_LINSEG_ONE=1 turns them on there (a tuning
// switch).  Smooth far blocks (no test at all) run in every mode.
#ifndef FJ_FAR_LINSEG_ONE
#define FJ_FAR_LINSEG_ONE 0
#endif
// Smooth far blocks in the one-spectrum modes pay in the dense kernel (J:
// 19,900 pairs x 4096 f x 3000 c in 248 ms against 300 ms without) and not
// in the kernels for small arrays (SMALL: the persistent batched kernel and
// the item-list kernel), whose blocks are seldom smooth: 4096 ragged
// sub-arrays 602 ms without against 635 ms with, the learned bands of a
// 300-pair field 8.9 ms against 9.6 ms (-DFJ_SMOOTH_SMALL=1 turns them on,
// a tuning switch).
#ifndef FJ_SMOOTH_SMALL
#define FJ_SMOOTH_SMALL 0
#endif
template <int MODE, bool SMALL = false>
struct LinFar {
  static constexpr bool on = ModeT<MODE>::nspec == 2 || FJ_FAR_LINSEG_ONE != 0;
  static constexpr bool smooth = ModeT<MODE>::nspec == 2 || !SMALL || FJ_SMOOTH_SMALL != 0;
};

// Stations per block of the linear-segment loop: 4, not kUnroll.  With 8,
// the records and the station data exceed the scalar registers and spill to
// vector lanes (J - Y: 718 ms against 447 ms per 4096 dense columns; J:
// 280 ms against 249 ms; 4096 ragged sub-arrays 638 ms against 602 ms).
// The block words are laid out for blocks of 4.  -DFJ_PRE_SMALL=0 turns the
// b prefetch off in the small-array kernels (646 ms against 631 ms), a
// tuning switch.
constexpr int kLinBlock = 4;
#ifndef FJ_PRE_SMALL
#define FJ_PRE_SMALL 1
#endif
// cf: the coefficients of one frequency (coef_block floats): the hat
// records (4 nspec floats per station), the b plane (nspec per station),
// the kappa plane (nrp per spectrum), then the long-segment limit
// (lin_header).  The lane sums the near stations' three kinds of term and
// the far stations' four in float partial sums and, every kPartial
// stations, adds them with their scales to a double accumulator.
template <int MODE, bool SMALL = false>
FJ_HD inline float linseg_lane(const Stations& S, const float* __restrict__ cf, int nrp, double q) {
  static_assert(kUnroll % kLinBlock == 0 && kPartial % kLinBlock == 0 && kLinBlock == 4,
                "station block");
  using LF = LinFar<MODE, SMALL>;
  using M = ModeT<MODE>;
  constexpr int NS = M::nspec;
  constexpr int NH = 4 * NS;
  constexpr int LU = kLinBlock;
  const float* __restrict__ cfb = cf + lin_b_offset(nrp, NS);
  const float* __restrict__ cfk = cf + lin_kappa_offset(nrp, NS);
  const uint32_t hlim = f32_bits(cf[lin_header(nrp, NS)]);
  const uint32_t alim = f32_bits(cf[lin_header(nrp, NS) + 1]);
  const LaneK L = lane_consts(q);
  const uint32_t irfar = far_threshold(L.q2pi);
  const uint32_t hsm = smooth_threshold(L.q2pi, hlim);
  LinState<MODE> st;
  double accR = 0.0, accI = 0.0;
  const double k1 = L.q2pi, k2 = k1 * k1, k3 = k2 * k1;
  const double sk = sqrt(2.0 * L.q2pi / kPiD);   // k ak; 0 at k = 0
  auto hat = [&](float a, float t, float p, float qq) {
    return sk * (((double)a - k2 * (double)t) + k1 * ((double)p - k2 * (double)qq));
  };
  auto flush = [&](LinState<MODE>& t) {
    accR += (double)t.aR + k1 * (double)t.pR + k3 * (double)t.tR + hat(t.haR, t.htR, t.hpR, t.hqR);
    t.aR = t.pR = t.tR = t.haR = t.htR = t.hpR = t.hqR = 0.0f;
    if (MODE == MODE_ABS_J || MODE == MODE_ABS_H) {
      accI += (double)t.aI + k1 * (double)t.pI + k3 * (double)t.tI +
              hat(t.haI, t.htI, t.hpI, t.hqI);
      t.aI = t.pI = t.tI = t.haI = t.htI = t.hpI = t.hqI = 0.0f;
    }
  };
  for (int sb = 0; sb < nrp; sb += LU) {
    float iv[LU];
    float hv[LU * NH];
#pragma unroll
    for (int k = 0; k < LU; k++) iv[k] = S.ir[sb + k];
#pragma unroll
    for (int k = 0; k < LU * NH; k++) hv[k] = cf[sb * NH + k];
    // the block's four words in one scalar load: a test of one after
    // another would wait for memory twice per block (sb is a multiple of 4
    // and the array 16-byte aligned; written by another kernel, so the
    // wider read aliases no store of this one)
    uint32_t bwv[4];
#if defined(__HIP_DEVICE_COMPILE__)
    {
      const uint4 b4 = reinterpret_cast<const uint4*>(S.bw)[sb >> 2];
      bwv[0] = b4.x;
      bwv[1] = b4.y;
      bwv[3] = b4.w;
    }
#else
    memcpy(bwv, S.bw + sb, 16);
#endif
    const uint32_t bw0 = bwv[0], bw1 = bwv[1];
    float ph[LU];
    block_phases<LU>(S, sb, L, Anchored<MODE>::on && bwv[3] <= alim, ph);
    const bool far = (LF::smooth || LF::on) && bw0 <= irfar;
    if (LF::smooth && far && bw1 <= hsm) {
      // every gap of the block (its left one included) is below both the
      // G-difference and the long-segment limits on every lane
      float kv[M::Y ? LU * NS : 1];
      if (M::Y) {
#pragma unroll
        for (int m = 0; m < NS; m++)
#pragma unroll
          for (int k = 0; k < LU; k++) kv[k * NS + m] = cfk[(size_t)m * nrp + sb + k];
      }
      // the block's far units first (independent), then the sums in
      // station order: the same operations as linseg_station's smooth path
      FarUnits w[LU];
#pragma unroll
      for (int k = 0; k < LU; k++) w[k] = far_units<M::Y>(ph[k], L.zk * iv[k]);
#pragma unroll
      for (int k = 0; k < LU; k++) {
        far_acc<MODE>(hv + k * NH, w[k], st);
        if (M::Y) kappa_acc<MODE>(kv + k * NS, st.aR, st.aI);
      }
      // smooth: the block's last segment is neither long nor a G difference
      st.gnext = st.lnext = false;
    } else {
      double rv[LU];
      float hw[LU + 1], sv[LU];   // hw[k]: gap left of station sb + k
#pragma unroll
      for (int k = 0; k < LU; k++) rv[k] = S.rr[sb + k];
#pragma unroll
      for (int k = 0; k <= LU; k++) hw[k] = S.hh[sb + k];
#pragma unroll
      for (int k = 0; k < LU; k++) sv[k] = S.isr[sb + k];
      // One-spectrum modes fetch the block's b records here, once, rather
      // than by a wait in each station's rare branch (not rare in arrays
      // of long segments: batched sub-arrays 658 -> 643 ms).  The
      // two-spectrum modes read them in the branch: the 8 more scalar
      // registers spill (J - Y 408 -> 438 ms per 4096 dense columns).
      constexpr bool PRE = NS == 1 && (!SMALL || FJ_PRE_SMALL != 0);
      float bv[PRE ? LU * NS : 1];
      if (PRE) {
#pragma unroll
        for (int k = 0; k < LU * NS; k++) bv[k] = cfb[(size_t)sb * NS + k];
      }
      const float* bsrc = PRE ? bv : cfb + (size_t)sb * NS;
      auto brec = [&](int k) { return bsrc + k * NS; };
      if (LF::on && far) {
#pragma unroll
        for (int k = 0; k < LU; k++)
          linseg_station<MODE, true>(sb + k, rv[k], ph[k], iv[k], sv[k], L.kf * hw[k],
                                     L.kf * hw[k + 1], hw[k + 1], hlim, L, hv + k * NH, nullptr,
                                     brec(k), st);
      } else {
#pragma unroll
        for (int k = 0; k < LU; k++)
          linseg_station<MODE, false>(sb + k, rv[k], ph[k], iv[k], sv[k], L.kf * hw[k],
                                      L.kf * hw[k + 1], hw[k + 1], hlim, L, hv + k * NH,
                                      nullptr, brec(k), st);
      }
    }
    if (((sb + LU) & (kPartial - 1)) == 0) flush(st);
  }
  flush(st);
  const double inv_k2 = 1.0 / (L.q2pi * L.q2pi);
  if (MODE == MODE_ABS_J || MODE == MODE_ABS_H)
    return (float)(sqrt(accR * accR + accI * accI) * inv_k2);
  return (float)(accR * inv_k2);
}

}  // namespace fjl

// back to the build's default (contraction where the language allows it)
#if defined(__clang__)
#pragma clang fp contract(fast)
#endif
