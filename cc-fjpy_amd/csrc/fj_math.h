// fj_math.h -- Bessel J0/J1/Y0/Y1 and the integral G(x) = int_0^x t Z1(t) dt
// for the F-J kernels, written once for the GPU (gfx950) and the host.
//
// Precision plan (see the kernels for how the pieces are used):
//   * The phase of the Hankel asymptotic form is reduced in double by the
//     caller: u = (x / 2pi - 1/8) - rint(.), a revolution count in
//     [-0.5, 0.5].  Everything after the reduction is float.
//   * x >= 8: Hankel asymptotic form with fitted P_n, Q_n (fj_fits.h), one
//     sin/cos pair shared by J0, J1, Y0, Y1, and the smooth Struve-minus-
//     Neumann fits R0, R1 for G.
//   * x < 8: double-precision polynomials in y = (x/8)^2 (fj_fits.h).
//
// All functions are __host__ __device__ so the host test build exercises
// the same arithmetic.  On the device the hardware sin/cos/rsq/log
// instructions are used when FJ_HW_TRANSCENDENTALS is 1; the host build
// uses the portable polynomial/libm path.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FJ_HD __host__ __device__
#else
#define FJ_HD
#endif

#include "fj_fits.h"

#ifndef FJ_HW_TRANSCENDENTALS
#define FJ_HW_TRANSCENDENTALS 1
#endif

#if defined(__HIP_DEVICE_COMPILE__) && FJ_HW_TRANSCENDENTALS
#define FJ_USE_HW 1
#else
#define FJ_USE_HW 0
#endif

namespace fjm {

constexpr float kSplit = 8.0f;                       // small/large argument split
constexpr float kSqrt2OverPi = 0.7978845608028654f;  // sqrt(2/pi)
constexpr double kTwoPi = 6.283185307179586476925;
constexpr float kTwoOverPi = 0.6366197723675814f;
constexpr double kTwoOverPiD = 0.63661977236758134308;
constexpr float kLn2 = 0.6931471805599453f;
constexpr float kPiOver2 = 1.5707963267948966f;
constexpr double kPiD = 3.14159265358979323846;

// ---- elementary pieces -----------------------------------------------------

// sin(2 pi u), cos(2 pi u) for u in [-0.5, 0.5] (revolutions).
FJ_HD inline void sincos_rev(float u, float* s, float* c) {
#if FJ_USE_HW
  *s = __builtin_amdgcn_sinf(u);
  *c = __builtin_amdgcn_cosf(u);
#else
  // quadrant split, then Taylor polynomials on |a| <= pi/4
  // (truncation below 2e-9 for sin and 1e-10 for cos).
  float q = rintf(4.0f * u);
  float a = (u - 0.25f * q) * 6.2831853071795865f;
  float a2 = a * a;
  float sp = __builtin_fmaf(a2, -2.5052108385441720e-08f, 2.7557319223985893e-06f);
  sp = __builtin_fmaf(sp, a2, -1.9841269841269841e-04f);
  sp = __builtin_fmaf(sp, a2, 8.3333333333333333e-03f);
  sp = __builtin_fmaf(sp, a2, -1.6666666666666667e-01f);
  sp = __builtin_fmaf(sp * a2, a, a);
  float cp = __builtin_fmaf(a2, 2.0876756987868099e-09f, -2.7557319223985893e-07f);
  cp = __builtin_fmaf(cp, a2, 2.4801587301587302e-05f);
  cp = __builtin_fmaf(cp, a2, -1.3888888888888889e-03f);
  cp = __builtin_fmaf(cp, a2, 4.1666666666666667e-02f);
  cp = __builtin_fmaf(cp, a2, -0.5f);
  cp = __builtin_fmaf(cp, a2, 1.0f);
  int iq = ((int)q) & 3;
  float ss = sp, cc = cp;
  if (iq == 1) { ss = cp; cc = -sp; }
  else if (iq == 2) { ss = -sp; cc = -cp; }
  else if (iq == 3) { ss = -cp; cc = sp; }
  *s = ss;
  *c = cc;
#endif
}

FJ_HD inline float rsqrt_f(float x) {
#if FJ_USE_HW
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}

// square root, about 1 ulp (the hardware instruction, no IEEE refinement)
FJ_HD inline float sqrt_f(float x) {
#if FJ_USE_HW
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}

FJ_HD inline float rcp_f(float x) {
#if FJ_USE_HW
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// natural log, float accuracy
FJ_HD inline float ln_f(float x) {
#if FJ_USE_HW
  return __builtin_amdgcn_logf(x) * kLn2;
#else
  return logf(x);
#endif
}

// ---- Bessel functions --------------------------------------------------------

struct Bess {
  float j0, j1, y0, y1;
  float gy;  // x Y1(x) + 2/pi, formed without cancellation near x = 0
};

// order-0 only (trapezoid kernels)
template <bool NEED_J, bool NEED_Y>
FJ_HD inline void bessel0_large(float x, float u, float* j0, float* y0) {
  float rs = rsqrt_f(x);
  float z = 8.0f * rs * rs;
  float t = z * z;
  float amp = kSqrt2OverPi * rs;
  float s, c;
  sincos_rev(u, &s, &c);
  float p0 = fit_P0(t), q0 = z * fit_q0(t);
  if (NEED_J) *j0 = amp * __builtin_fmaf(p0, c, -q0 * s);
  if (NEED_Y) *y0 = amp * __builtin_fmaf(p0, s, q0 * c);
}

// x < 8 in double.  lnx2pi = (2/pi) ln x (float accuracy suffices: it only
// multiplies J, whose size is at most 1).  TAB picks how the polynomial
// coefficients are held (FIT_LITERAL / FIT_TABLE); the best choice depends
// on the calling kernel's register pressure.
template <bool NEED_J1, bool NEED_Y, int TAB = FIT_LITERAL>
FJ_HD inline Bess bessel_small(double x, double* y_out, double* lnterm_out) {
  double y = (x * 0.125) * (x * 0.125);
  Bess b;
  double j0 = fit_J0s<TAB>(y);
  double j1 = NEED_J1 ? x * fit_J1s<TAB>(y) : 0.0;
  b.j0 = (float)j0;
  b.j1 = (float)j1;
  b.y0 = b.y1 = b.gy = 0.0f;
  double L = 0.0;
  if (NEED_Y) {
    L = (double)(kTwoOverPi * ln_f((float)x));
    b.y0 = (float)(L * j0 + fit_Y0s<TAB>(y));
    if (NEED_J1) {
      const double xf1 = x * fit_Y1s<TAB>(y);
      const double inv = (double)(kTwoOverPi * rcp_f((float)x));
      b.y1 = (float)(L * j1 - inv + xf1);
      b.gy = (float)(x * (L * j1 + xf1));  // x Y1 + 2/pi
    }
  }
  *y_out = y;
  *lnterm_out = L;
  return b;
}

// G(x), small x (x < 8), from the double polynomials.  For the Y basis it
// returns the shifted integral G_Y(x) + (2/pi) x = int_0^x (t Y1 + 2/pi) dt,
// which vanishes like x^3 ln x instead of approaching -(2/pi) x.
template <int TAB = FIT_LITERAL>
FJ_HD inline void g_small(double x, double y, double L, bool need_j, bool need_y,
                          float* gj, float* gy_shifted) {
  double g = x * x * x * fit_GJs<TAB>(y);
  if (need_j) *gj = (float)g;
  // x ln x -> 0 at x = 0; avoid ln(0) * 0
  if (need_y)
    *gy_shifted = x > 0.0 ? (float)(L * g + x * (fit_GYs<TAB>(y) + kTwoOverPiD)) : 0.0f;
}

// cos(2 pi u) and sin(2 pi u) alone (revolutions; u within a few turns)
FJ_HD inline float cos_rev(float u) {
#if FJ_USE_HW
  return __builtin_amdgcn_cosf(u);
#else
  float s, c;
  sincos_rev(u, &s, &c);
  return c;
#endif
}

FJ_HD inline float sin_rev(float u) {
#if FJ_USE_HW
  return __builtin_amdgcn_sinf(u);
#else
  float s, c;
  sincos_rev(u, &s, &c);
  return s;
#endif
}

// Reduced phase of x - pi/4 in revolutions: q = f/c (cycles per metre),
// r in metres, x / 2pi = q r.
FJ_HD inline float reduce_phase(double q, double r) {
  double t = __builtin_fma(q, r, -0.125);
  return (float)(t - rint(t));
}

// t - floor(t) for finite t (one v_fract_f64 on the device).
FJ_HD inline double fract_d(double t) {
#if FJ_USE_HW
  return __builtin_amdgcn_fract(t);
#else
  return t - floor(t);
#endif
}

// The same phase reduced to [0, 1) (one v_fract_f64 on the device).  The
// float rounding of the result is at most 2^-25 revolutions, as for
// reduce_phase, and sincos_rev accepts either range.
FJ_HD inline float phase_rev(double q, double r) {
  return (float)fract_d(__builtin_fma(q, r, -0.125));
}

}  // namespace fjm
