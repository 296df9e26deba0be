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

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/ccfj.h"
#include "fj_math.h"

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
};

// Entry s of the station arrays from the sorted float distances r[0..nr).
FJ_HD inline void station_entry(const float* r, int nr, int s, double* rr, float* hh, float* ir,
                                float* isr) {
  const bool in = s < nr;
  const double rs = (double)r[in ? s : nr - 1];
  rr[s] = rs;
  hh[s] = (in && s > 0) ? r[s] - r[s - 1] : 0.0f;
  ir[s] = (float)(1.0 / rs);
  isr[s] = (float)(1.0 / sqrt(rs));
}

// Per-lane constants of one (f, c).
struct LaneK {
  double q;     // f / c, cycles per metre
  double q2pi;  // k = 2 pi f / c
  float kf;     // k in float
  float zk;     // 8 / k           (z = 8/x = zk / r)
  float ak;     // sqrt(2/(pi k))   (Hankel amplitude = ak / sqrt r)
};

FJ_HD inline LaneK lane_consts(double q) {
  LaneK L;
  L.q = q;
  L.q2pi = q * kTwoPi;
  L.kf = (float)L.q2pi;
  L.zk = (float)(8.0 / L.q2pi);
  L.ak = (float)sqrt(2.0 / (kPiD * L.q2pi));
  return L;
}

// ---- wave-uniform far blocks -----------------------------------------------
// A block of kUnroll stations is "far" when x >= kFarX for every lane of the
// wave, i.e. when every station of the block has 1/r <= kmin / kFarX with
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
  return __builtin_amdgcn_readfirstlane(f32_bits((float)(kmin / (double)kFarX)));
#else
  return f32_bits((float)(kmin / (double)kFarX));
#endif
}

template <int N>
FJ_HD inline bool block_far(const float* iv, uint32_t irfar) {
  uint32_t m = f32_bits(iv[0]);
#pragma unroll
  for (int k = 1; k < N; k++) {
    const uint32_t b = f32_bits(iv[k]);
    m = b > m ? b : m;
  }
  return m <= irfar;
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

FJ_HD inline int coef_count(int mode, int itype) {
  int nspec = (mode == MODE_J || mode == MODE_Y) ? 1 : 2;
  return itype == 0 ? nspec : 2 * nspec;
}

// Occupancy floor of the F-J kernels (waves per SIMD; the compiler then
// keeps VGPRs within 512 / waves), used by the dense, batched and item-list
// kernels alike.  Measured on the MI355X: the fused J - Y
// linear-segment kernel (natural 217 VGPRs, 2 waves) runs 16% faster held
// to 3 waves, a few spills on its cold small-argument path included; the
// other modes are fastest at their natural count (forcing them spills on the
// hot path).  The linear-segment J kernel holds 4 waves (121 VGPRs in the
// dense kernel; the persistent batched kernel's loop would take 170).
// -DFJ_OCC_I1_JMY=N / -DFJ_OCC_I1_J=N override for tuning runs.
#ifndef FJ_OCC_I1_JMY
#define FJ_OCC_I1_JMY 3
#endif
#ifndef FJ_OCC_I1_J
#define FJ_OCC_I1_J 4
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
template <int MODE, int ITYPE>
struct FjOcc {
  static constexpr int waves =
      ITYPE == 0 ? 1 : (MODE == MODE_J_MINUS_Y ? FJ_OCC_I1_JMY : (MODE == MODE_J ? FJ_OCC_I1_J : 1));
};

// ---- coefficients of station s (one frequency) ----------------------------
// g_prev/g_cur: spectrum values at stations s-1 and s (g_prev unused for
// s = 0).  rp, rs, rn: r_{s-1}, r_s, r_{s+1} (only those that exist are read).
// itype 0 writes A_s per spectrum; itype 1 writes (alpha_s, beta_s).
template <int ITYPE, int NSPEC>
FJ_HD inline void station_coef(int s, int nr, float rp, float rs, float rn,
                               const float* g_prev, const float* g_cur, float* rec) {
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
    // dr0 = max(h, 0.1); the segment's line reaches g1 + b h at r_s.
    double ratio = 0.0;
    if (s > 0) {
      const float h = rs - rp;
      const float dr0 = fmaxf(h, 0.1f);
      ratio = (double)h / (double)dr0;
    }
    for (int k = 0; k < NSPEC; k++) {
      const double g2 = g_cur[k];
      const double g1 = s > 0 ? (double)g_prev[k] : 0.0;
      const double beta = s > 0 ? (g2 - g1) * ratio : 0.0;
      const double el = s > 0 ? g1 + beta : 0.0;  // left limit at r_s
      const double er = s + 1 < nr ? g2 : 0.0;    // right limit at r_s
      rec[2 * k] = (float)(el - er);              // alpha_s
      rec[2 * k + 1] = (float)(0.5 * beta);       // beta_s / 2 (the lane forms 2 D_s)
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
template <int MODE>
FJ_HD inline void trapz_station(double rs, float ir, const LaneK& L, const float* c, float& pR,
                                float& pI) {
  using M = ModeT<MODE>;
  float vj = 0.0f, vy = 0.0f;
  const float z = L.zk * ir;  // 8 / x; inf when k = 0 or r = 0
  if (z <= 1.0f) {            // x >= 8
    const float t = z * z;
    float p0, q0;
    if (z <= 8.0f / kFarX) {
      p0 = fitf_P0(t);
      q0 = z * fitf_q0(t);
    } else {
      p0 = fit_P0(t);
      q0 = z * fit_q0(t);
    }
    float s, co;
    sincos_rev(phase_rev(L.q, rs), &s, &co);
    if (M::J) vj = __builtin_fmaf(p0, co, -q0 * s);
    if (M::Y) vy = __builtin_fmaf(p0, s, q0 * co);
  } else {
    const double xd = L.q2pi * rs;
    double y, Lg;
    const Bess b = bessel_small<false, M::Y, FIT_TABLE>(xd, &y, &Lg);
    const float ia = L.kf > 0.0f ? sqrtf(kPiOver2 * (float)xd) : sqrtf((float)rs);
    vj = b.j0 * ia;
    vy = b.y0 * ia;
  }
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

// x >= kFarX on every lane (a far block): no branch on the argument.
template <int MODE>
FJ_HD inline void trapz_far(double rs, float ir, const LaneK& L, const float* c, float& pR,
                            float& pI) {
  using M = ModeT<MODE>;
  // modulus-phase form: J0 = amp M0 cos(th + phi), Y0 = amp M0 sin(th + phi),
  // one transcendental per basis (the amplitude is applied per lane)
  const float z = L.zk * ir, t = z * z;
  const float m0 = fitf_M0(t);
  const float u = __builtin_fmaf(z, fitf_ph0(t), phase_rev(L.q, rs));
  float vj = 0.0f, vy = 0.0f;
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
  } else {
    pR = __builtin_fmaf(c[0], vj, pR);
    pR = __builtin_fmaf(-c[1], vy, pR);
    pI = __builtin_fmaf(c[1], vj, pI);
    pI = __builtin_fmaf(c[0], vy, pI);
  }
}

template <int MODE>
FJ_HD inline float trapz_lane(const Stations& S, const float* __restrict__ cf, int nrp, double q) {
  constexpr int NC = ModeT<MODE>::nspec;
  const LaneK L = lane_consts(q);
  const uint32_t irfar = far_threshold(L.q2pi);
  double accR = 0.0, accI = 0.0;
  for (int s0 = 0; s0 < nrp; s0 += kPartial) {
    const int s1 = nrp < s0 + kPartial ? nrp : s0 + kPartial;
    float pR = 0.0f, pI = 0.0f;
    for (int sb = s0; sb < s1; sb += kUnroll) {
      double rv[kUnroll];
      float iv[kUnroll];
      float cv[kUnroll * NC];
#pragma unroll
      for (int k = 0; k < kUnroll; k++) rv[k] = S.rr[sb + k];
#pragma unroll
      for (int k = 0; k < kUnroll; k++) iv[k] = S.ir[sb + k];
#pragma unroll
      for (int k = 0; k < kUnroll * NC; k++) cv[k] = cf[sb * NC + k];
      if (block_far<kUnroll>(iv, irfar)) {
#pragma unroll
        for (int k = 0; k < kUnroll; k++) trapz_far<MODE>(rv[k], iv[k], L, cv + k * NC, pR, pI);
      } else {
#pragma unroll
        for (int k = 0; k < kUnroll; k++)
          trapz_station<MODE>(rv[k], iv[k], L, cv + k * NC, pR, pI);
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
// The lane forms E = 2 D (and 2 D~) and the coefficients carry beta / 2.
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

// x >= 8 at station (rs, ir = 1/r, isr = 1/sqrt r) for lane L: J0, J1,
// (NY) Y0, Y1, and (need_g) G_J = 1 + R0 J1 - x R1 J0, G_Y = R0 Y1 - x R1 Y0.
// Degree-2/3 fits from x = 20 on, degree 4-7 below; FAR_ONLY: the caller
// knows x >= 20 on every lane.
template <bool NJ, bool NY, bool FAR_ONLY = false>
FJ_HD inline void large_station(const LaneK& L, double rs, float ir, float isr, float x,
                                bool need_g, float* j0, float* j1, float* y0, float* y1,
                                float* GJ, float* GY) {
  const float z = L.zk * ir, t = z * z;
  const float amp = L.ak * isr, az = amp * z;
  const bool far = FAR_ONLY || x >= kFarX;
  float P0, Q0, P1, Q1;
  if (far) {
    P0 = fitf_P0(t);
    Q0 = fitf_q0(t);
    P1 = fitf_P1(t);
    Q1 = fitf_q1(t);
  } else {
    P0 = fit_P0(t);
    Q0 = fit_q0(t);
    P1 = fit_P1(t);
    Q1 = fit_q1(t);
  }
  float sn, co;
  sincos_rev(phase_rev(L.q, rs), &sn, &co);
  const float A = amp * P0, B = az * Q0, C = amp * P1, D = az * Q1;
  *j0 = __builtin_fmaf(A, co, -B * sn);
  *j1 = __builtin_fmaf(C, sn, D * co);
  if (NY) {
    *y0 = __builtin_fmaf(A, sn, B * co);
    *y1 = __builtin_fmaf(D, sn, -C * co);
  }
  if (need_g) {
    float r0, r1;
    if (far) {
      r0 = fitf_R0(t);
      r1 = fitf_R1(t);
    } else {
      r0 = fit_R0(t);
      r1 = fit_R1(t);
    }
    const float xr1 = x * r1;
    if (NJ) *GJ = __builtin_fmaf(r0, *j1, __builtin_fmaf(-xr1, *j0, 1.0f));
    if (NY) *GY = __builtin_fmaf(r0, *y1, -xr1 * *y0);
  }
}

template <int MODE>
struct LinState {
  Ends pj = {0.f, 0.f, 0.f, 0.f, 0.f}, py = {0.f, 0.f, 0.f, 0.f, 0.f};
  float xprev = 0.0f;
  bool gnext = false;  // the segment right of the previous station uses the G difference
  float pR = 0.0f, pI = 0.0f;
};

// End of a station: its g, g' and g''' per basis, the divided difference
// of the segment to its left (2 D_s: Euler-Maclaurin, or the G difference
// on the lanes whose segment needs it -- evaluated only when some lane of
// the wave does), and the station's contribution alpha*g - (beta/2)*2D.
template <int MODE, bool FAR>
FJ_HD inline void linseg_tail(int s, float x, float dxL, bool useG_L, float j0, float j1, float y0,
                              float y1, float gy, Ends& cj, Ends& cy, const float* c,
                              LinState<MODE>& st) {
  using M = ModeT<MODE>;
  constexpr bool NJ = M::J, NY = M::Y;
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
  float Ej = 0.0f, Ey = 0.0f;  // 2 D_s per basis
  // Two forms, chosen by measurement on the MI355X: straight-line (far
  // blocks, and the two-spectrum modes everywhere) or a per-lane branch (the
  // one-spectrum modes' general loop, which it keeps lighter in registers).
  constexpr bool SELECT = FAR || ModeT<MODE>::nspec == 2;
  if (SELECT && s > 0) {
    // the rule everywhere, the G difference selected on the lanes that need
    // it when any does (a wave-uniform branch)
    const float e1 = dxL * (-1.0f / 6.0f);
    const float e3 = dxL * dxL * dxL * (1.0f / 360.0f);
    if (NJ) Ej = em_divided2(st.pj, cj, e1, e3);
    if (NY) Ey = em_divided2(st.py, cy, e1, e3);
    if (wave_any(useG_L)) {
      const float inv2 = rcp_f(0.5f * dxL);
      if (NJ) {
        const float g = (cj.G - st.pj.G) * inv2;
        Ej = useG_L ? g : Ej;
      }
      if (NY) {
        const float g = gdiff_y2(st.py, st.xprev, cy, x, inv2);
        Ey = useG_L ? g : Ey;
      }
    }
  } else if (s > 0) {
    // per lane: one of the two (fewer live registers)
    if (useG_L) {
      const float inv2 = rcp_f(0.5f * dxL);
      if (NJ) Ej = (cj.G - st.pj.G) * inv2;
      if (NY) Ey = gdiff_y2(st.py, st.xprev, cy, x, inv2);
    } else {
      const float e1 = dxL * (-1.0f / 6.0f);
      const float e3 = dxL * dxL * dxL * (1.0f / 360.0f);
      if (NJ) Ej = em_divided2(st.pj, cj, e1, e3);
      if (NY) Ey = em_divided2(st.py, cy, e1, e3);
    }
  }
  st.xprev = x;
  if (MODE == MODE_J) {
    st.pR = __builtin_fmaf(c[0], cj.g, st.pR);
    st.pR = __builtin_fmaf(-c[1], Ej, st.pR);
  } else if (MODE == MODE_Y) {
    st.pR = __builtin_fmaf(c[0], cy.g, st.pR);
    st.pR = __builtin_fmaf(-c[1], Ey, st.pR);
  } else if (MODE == MODE_J_MINUS_Y) {
    st.pR = __builtin_fmaf(c[0], cj.g, st.pR);
    st.pR = __builtin_fmaf(-c[1], Ej, st.pR);
    st.pR = __builtin_fmaf(-c[2], cy.g, st.pR);
    st.pR = __builtin_fmaf(c[3], Ey, st.pR);
  } else if (MODE == MODE_ABS_J) {
    st.pR = __builtin_fmaf(c[0], cj.g, st.pR);
    st.pR = __builtin_fmaf(-c[1], Ej, st.pR);
    st.pI = __builtin_fmaf(c[2], cj.g, st.pI);
    st.pI = __builtin_fmaf(-c[3], Ej, st.pI);
  } else {  // MODE_ABS_H
    const float jre = __builtin_fmaf(c[0], cj.g, -c[1] * Ej);
    const float jim = __builtin_fmaf(c[2], cj.g, -c[3] * Ej);
    const float yre = __builtin_fmaf(c[0], cy.g, -c[1] * Ey);
    const float yim = __builtin_fmaf(c[2], cy.g, -c[3] * Ey);
    st.pR += jre - yim;
    st.pI += jim + yre;
  }
  st.pj = cj;
  st.py = cy;
}

// One station of the linear-segment sum.  dxL = k (r_s - r_{s-1}),
// dxR = k (r_{s+1} - r_s) (0 past the last station), c = alpha, beta/2
// record per spectrum.  FAR: x >= kFarX on every lane of the wave.
template <int MODE, bool FAR>
FJ_HD inline void linseg_station(int s, double rs, float ir, float isr, float dxL, float dxR,
                                 const LaneK& L, const float* c, LinState<MODE>& st) {
  using M = ModeT<MODE>;
  constexpr bool NJ = M::J, NY = M::Y;
  const double xd = L.q2pi * rs;
  const float x = (float)xd;
  const bool useG_L = s > 0 && st.gnext;
  // in a far block both ends of every segment have x >= kFarX, where the
  // rule's threshold min(kTau, kKappa min(x_a, x_b)) is kTau
  const bool gR = FAR ? fabsf(dxR) >= kTau : use_gdiff_right(x, dxR);
  st.gnext = gR;
  const bool needG = useG_L || gR;
  Ends cj = {0.f, 0.f, 0.f, 0.f, 0.f}, cy = {0.f, 0.f, 0.f, 0.f, 0.f};
  float j0, j1, y0 = 0.0f, y1 = 0.0f, gy = 0.0f;
  if (FAR) {
    large_station<NJ, NY, true>(L, rs, ir, isr, x, wave_any(needG), &j0, &j1, &y0, &y1, &cj.G,
                                &cy.G);
    if (NY) gy = __builtin_fmaf(x, y1, kTwoOverPi);
  } else if (x >= kSplit) {
    large_station<NJ, NY>(L, rs, ir, isr, x, needG, &j0, &j1, &y0, &y1, &cj.G, &cy.G);
    if (NY) gy = __builtin_fmaf(x, y1, kTwoOverPi);
  } else {
    double y, Lg;
    const Bess b = bessel_small<true, NY, FJ_FIT_I1>(xd, &y, &Lg);
    j0 = b.j0;
    j1 = b.j1;
    y0 = b.y0;
    y1 = b.y1;
    gy = b.gy;
    if (needG) g_small<FJ_FIT_I1>(xd, y, Lg, NJ, NY, &cj.G, &cy.G);
    cy.sh = 1.0f;
  }
  linseg_tail<MODE, FAR>(s, x, dxL, useG_L, j0, j1, y0, y1, gy, cj, cy, c, st);
}

// Far blocks in the linear-segment loop.  Measured on the MI355X they pay
// for the two-spectrum modes (J - Y: 69 -> 58 ms per 512 dense columns) and
// not for the one-spectrum modes, whose general loop is lighter in
// registers; -DFJ_FAR_LINSEG_ONE=1 turns them on there (a tuning switch).
#ifndef FJ_FAR_LINSEG_ONE
#define FJ_FAR_LINSEG_ONE 0
#endif
template <int MODE>
struct LinFar {
  static constexpr bool on = ModeT<MODE>::nspec == 2 || FJ_FAR_LINSEG_ONE != 0;
};

// cf: NC = 2 nspec floats per station: (alpha, beta/2) per spectrum.
template <int MODE>
FJ_HD inline float linseg_lane(const Stations& S, const float* __restrict__ cf, int nrp, double q) {
  constexpr int NC = 2 * ModeT<MODE>::nspec;
  const LaneK L = lane_consts(q);
  const uint32_t irfar = far_threshold(L.q2pi);
  LinState<MODE> st;
  double accR = 0.0, accI = 0.0;
  float dxR = 0.0f;
  for (int sb = 0; sb < nrp; sb += kUnroll) {
    double rv[kUnroll];
    float hv[kUnroll], iv[kUnroll + 1], sv[kUnroll];
    float cv[kUnroll * NC];
#pragma unroll
    for (int k = 0; k < kUnroll; k++) rv[k] = S.rr[sb + k];
#pragma unroll
    for (int k = 0; k < kUnroll; k++) hv[k] = S.hh[sb + k + 1];
#pragma unroll
    for (int k = 0; k <= kUnroll; k++) iv[k] = S.ir[sb + k];   // + the next block's first
#pragma unroll
    for (int k = 0; k < kUnroll; k++) sv[k] = S.isr[sb + k];
#pragma unroll
    for (int k = 0; k < kUnroll * NC; k++) cv[k] = cf[sb * NC + k];
    if (LinFar<MODE>::on && block_far<kUnroll + 1>(iv, irfar)) {
#pragma unroll
      for (int k = 0; k < kUnroll; k++) {
        const float dxL = dxR;
        dxR = L.kf * hv[k];
        linseg_station<MODE, true>(sb + k, rv[k], iv[k], sv[k], dxL, dxR, L, cv + k * NC, st);
      }
    } else {
#pragma unroll
      for (int k = 0; k < kUnroll; k++) {
        const float dxL = dxR;
        dxR = L.kf * hv[k];
        linseg_station<MODE, false>(sb + k, rv[k], iv[k], sv[k], dxL, dxR, L, cv + k * NC, st);
      }
    }
    if (((sb + kUnroll) & (kPartial - 1)) == 0) {
      accR += (double)st.pR;
      accI += (double)st.pI;
      st.pR = st.pI = 0.0f;
    }
  }
  accR += (double)st.pR;
  accI += (double)st.pI;
  const double inv_k2 = 1.0 / (L.q2pi * L.q2pi);
  if (MODE == MODE_ABS_J || MODE == MODE_ABS_H)
    return (float)(sqrt(accR * accR + accI * accI) * inv_k2);
  return (float)(accR * inv_k2);
}

}  // namespace fjl
