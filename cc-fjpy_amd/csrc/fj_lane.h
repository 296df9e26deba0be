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
#pragma once

#include <math.h>
#include <stdint.h>

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

// ---- coefficients of station s (one frequency) ----------------------------
// g_prev/g_cur: spectrum values at stations s-1 and s (g_prev unused for
// s = 0).  rp, rs, rn: r_{s-1}, r_s, r_{s+1} (only those that exist are read).
// itype 0 writes A_s per spectrum; itype 1 writes (alpha_s, beta_s).
template <int ITYPE, int NSPEC>
FJ_HD inline void station_coef(int s, int nr, float rp, float rs, float rn,
                               const float* g_prev, const float* g_cur, float* rec) {
  if (ITYPE == 0) {
    // trapezoid weight (src/FJcpu.cpp:354-356): each segment gives d/2 to
    // both of its end stations, d = max(h, 0.1) in float as the reference
    double w = 0.0;
    if (s > 0) w += (double)fmaxf(rs - rp, 0.1f);
    if (s + 1 < nr) w += (double)fmaxf(rn - rs, 0.1f);
    const double wr = 0.5 * w * (double)rs;
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
      rec[2 * k + 1] = (float)beta;               // beta_s
    }
  }
}

// ---- itype 0 ---------------------------------------------------------------
// q = f/c in cycles per metre (double).  cf: coefficient records of one
// frequency, NC = nspec floats per station; nrp = padded station count.
template <int MODE>
FJ_HD inline void trapz_station(double rs, double q, double q2pi, const float* c,
                                float& pR, float& pI) {
  using M = ModeT<MODE>;
  const double xd = q2pi * rs;
  const float x = (float)xd;
  float J0 = 0.0f, Y0 = 0.0f;
  if (x >= kSplit) {
    bessel0_large<M::J, M::Y>(x, reduce_phase(q, rs), &J0, &Y0);
  } else {
    double y, L;
    Bess b = bessel_small<false, M::Y>(xd, &y, &L);
    J0 = b.j0;
    Y0 = b.y0;
  }
  if (MODE == MODE_J) {
    pR = __builtin_fmaf(c[0], J0, pR);
  } else if (MODE == MODE_Y) {
    pR = __builtin_fmaf(c[0], Y0, pR);
  } else if (MODE == MODE_J_MINUS_Y) {
    pR = __builtin_fmaf(c[0], J0, pR);
    pR = __builtin_fmaf(-c[1], Y0, pR);
  } else if (MODE == MODE_ABS_J) {
    pR = __builtin_fmaf(c[0], J0, pR);
    pI = __builtin_fmaf(c[1], J0, pI);
  } else {  // MODE_ABS_H: (FJ[re]-FY[im]) + i (FJ[im]+FY[re])
    pR = __builtin_fmaf(c[0], J0, pR);
    pR = __builtin_fmaf(-c[1], Y0, pR);
    pI = __builtin_fmaf(c[1], J0, pI);
    pI = __builtin_fmaf(c[0], Y0, pI);
  }
}

template <int MODE>
FJ_HD inline float trapz_lane(const double* __restrict__ rr, const float* __restrict__ cf,
                              int nrp, double q) {
  constexpr int NC = ModeT<MODE>::nspec;
  const double q2pi = q * kTwoPi;
  double accR = 0.0, accI = 0.0;
  for (int s0 = 0; s0 < nrp; s0 += kPartial) {
    const int s1 = nrp < s0 + kPartial ? nrp : s0 + kPartial;
    float pR = 0.0f, pI = 0.0f;
    for (int sb = s0; sb < s1; sb += kUnroll) {
      double rv[kUnroll];
      float cv[kUnroll * NC];
#pragma unroll
      for (int k = 0; k < kUnroll; k++) rv[k] = rr[sb + k];
#pragma unroll
      for (int k = 0; k < kUnroll * NC; k++) cv[k] = cf[sb * NC + k];
#pragma unroll
      for (int k = 0; k < kUnroll; k++) trapz_station<MODE>(rv[k], q, q2pi, cv + k * NC, pR, pI);
    }
    accR += (double)pR;
    accI += (double)pI;
  }
  if (MODE == MODE_ABS_J || MODE == MODE_ABS_H) return (float)sqrt(accR * accR + accI * accI);
  return (float)accR;
}

// ---- itype 1 ---------------------------------------------------------------
// Y basis: x Y1(x) -> -2/pi as x -> 0.  Because sum_s alpha_s = sum_s beta_s
// exactly (the jumps and rises of one interpolant telescope), that constant
// drops out of alpha g - beta D; the Y terms therefore use the shifted
// g~ = x Y1 + 2/pi and D~ = D + 2/pi, which removes a cancellation of up to
// ~1/(x^2 ln x) at small x.  G is stored shifted (G + 2x/pi) where it comes
// from the small-argument branch (sh = 1) and unshifted otherwise.
struct Ends {  // per-basis values at a station
  float g, gp, g3, G, sh;
};

// (1/dx) int_{x_a}^{x_b} g by Euler-Maclaurin with two end corrections
FJ_HD inline float em_divided(const Ends& a, const Ends& b, float dx) {
  const float dx2 = dx * dx;
  float d = 0.5f * (a.g + b.g);
  d = __builtin_fmaf(-dx * (1.0f / 12.0f), b.gp - a.gp, d);
  d = __builtin_fmaf(dx * dx2 * (1.0f / 720.0f), b.g3 - a.g3, d);
  return d;
}

// D~ for the Y basis from stored (possibly shifted) G values.
FJ_HD inline float gdiff_y(const Ends& a, float xa, const Ends& b, float xb, float inv) {
  if (a.sh == b.sh) return __builtin_fmaf(b.G - a.G, inv, a.sh != 0.0f ? 0.0f : kTwoOverPi);
  const float Ga = a.sh != 0.0f ? a.G : __builtin_fmaf(kTwoOverPi, xa, a.G);
  const float Gb = b.sh != 0.0f ? b.G : __builtin_fmaf(kTwoOverPi, xb, b.G);
  return (Gb - Ga) * inv;
}

// cf: NC = 2 nspec floats per station: (alpha, beta) per spectrum.
template <int MODE>
struct LinState {
  Ends pj = {0.f, 0.f, 0.f, 0.f, 0.f}, py = {0.f, 0.f, 0.f, 0.f, 0.f};
  float xprev = 0.0f;
  float pR = 0.0f, pI = 0.0f;
};

// One station of the linear-segment sum.  dxL = k (r_s - r_{s-1}),
// dxR = k (r_{s+1} - r_s) (0 past the last station), c = alpha/beta record.
template <int MODE>
FJ_HD inline void linseg_station(int s, double rs, float dxL, float dxR, double q, double q2pi,
                                 const float* c, LinState<MODE>& st) {
  using M = ModeT<MODE>;
  constexpr bool NJ = M::J, NY = M::Y;
  const double xd = q2pi * rs;
  const float x = (float)xd;
  const bool useG_L = s > 0 && use_gdiff(st.xprev, x, dxL);
  const bool needG = useG_L || use_gdiff(x, x + dxR, dxR);
  Ends cj = {0.f, 0.f, 0.f, 0.f, 0.f}, cy = {0.f, 0.f, 0.f, 0.f, 0.f};
  Bess b;
  if (x >= kSplit) {
    float t;
    b = bessel_large<NJ, NY>(x, reduce_phase(q, rs), &t);
    if (needG) g_large(x, t, b, &cj.G, &cy.G, NJ, NY);
  } else {
    double y, L;
    b = bessel_small<true, NY>(xd, &y, &L);
    if (needG) g_small(xd, y, L, NJ, NY, &cj.G, &cy.G);
    cy.sh = 1.0f;
  }
  if (NJ) {
    cj.g = x * b.j1;
    cj.gp = x * b.j0;
    cj.g3 = -b.j1 - cj.gp;
  }
  if (NY) {
    cy.g = b.gy;
    cy.gp = x * b.y0;
    cy.g3 = -b.y1 - cy.gp;
  }
  float Dj = 0.0f, Dy = 0.0f;
  if (s > 0) {
    if (useG_L) {
      const float inv = rcp_f(dxL);
      if (NJ) Dj = (cj.G - st.pj.G) * inv;
      if (NY) Dy = gdiff_y(st.py, st.xprev, cy, x, inv);
    } else {
      if (NJ) Dj = em_divided(st.pj, cj, dxL);
      if (NY) Dy = em_divided(st.py, cy, dxL);
    }
  }
  st.xprev = x;
  // station s contributes alpha*g - beta*D per spectrum
  if (MODE == MODE_J) {
    st.pR += c[0] * cj.g - c[1] * Dj;
  } else if (MODE == MODE_Y) {
    st.pR += c[0] * cy.g - c[1] * Dy;
  } else if (MODE == MODE_J_MINUS_Y) {
    st.pR += (c[0] * cj.g - c[1] * Dj) - (c[2] * cy.g - c[3] * Dy);
  } else if (MODE == MODE_ABS_J) {
    st.pR += c[0] * cj.g - c[1] * Dj;
    st.pI += c[2] * cj.g - c[3] * Dj;
  } else {  // MODE_ABS_H
    const float jre = c[0] * cj.g - c[1] * Dj;
    const float jim = c[2] * cj.g - c[3] * Dj;
    const float yre = c[0] * cy.g - c[1] * Dy;
    const float yim = c[2] * cy.g - c[3] * Dy;
    st.pR += jre - yim;
    st.pI += jim + yre;
  }
  st.pj = cj;
  st.py = cy;
}

// cf: NC = 2 nspec floats per station: (alpha, beta) per spectrum.
// hh[s] = r_s - r_{s-1}, hh[0] = 0, zero past the last station; it holds
// nrp + kUnroll entries.
template <int MODE>
FJ_HD inline float linseg_lane(const double* __restrict__ rr, const float* __restrict__ hh,
                               const float* __restrict__ cf, int nrp, double q) {
  constexpr int NC = 2 * ModeT<MODE>::nspec;
  const double q2pi = q * kTwoPi;
  const float kf = (float)q2pi;
  LinState<MODE> st;
  double accR = 0.0, accI = 0.0;
  float dxR = 0.0f;
  for (int sb = 0; sb < nrp; sb += kUnroll) {
    double rv[kUnroll];
    float hv[kUnroll];
    float cv[kUnroll * NC];
#pragma unroll
    for (int k = 0; k < kUnroll; k++) rv[k] = rr[sb + k];
#pragma unroll
    for (int k = 0; k < kUnroll; k++) hv[k] = hh[sb + k + 1];
#pragma unroll
    for (int k = 0; k < kUnroll * NC; k++) cv[k] = cf[sb * NC + k];
#pragma unroll
    for (int k = 0; k < kUnroll; k++) {
      const float dxL = dxR;
      dxR = kf * hv[k];
      linseg_station<MODE>(sb + k, rv[k], dxL, dxR, q, q2pi, cv + k * NC, st);
    }
    if (((sb + kUnroll) & (kPartial - 1)) == 0) {
      accR += (double)st.pR;
      accI += (double)st.pI;
      st.pR = st.pI = 0.0f;
    }
  }
  accR += (double)st.pR;
  accI += (double)st.pI;
  const double inv_k2 = 1.0 / (q2pi * q2pi);
  if (MODE == MODE_ABS_J || MODE == MODE_ABS_H)
    return (float)(sqrt(accR * accR + accI * accI) * inv_k2);
  return (float)(accR * inv_k2);
}

}  // namespace fjl
