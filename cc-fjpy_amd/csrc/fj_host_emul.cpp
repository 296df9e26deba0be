// fj_host_emul.cpp -- host (CPU) build of the F-J kernel arithmetic.
//
// Compiles fj_lane.h with the portable sin/cos/rsq/log path so the CPU test
// suite can check the kernel algorithm (coefficient preparation, the
// Euler-Maclaurin / G-difference segment rule, the small/large argument
// split, compensated accumulation) against the FP64 oracle without a GPU.
// It is a test build of the kernel code, not a fallback: the ccfj Python
// package never loads it.
#include <stdint.h>
#include <stdlib.h>

#include <vector>

#include "fj_lane.h"

using namespace fjl;

namespace {

template <int MODE>
void run_mode(int itype, const float* U0, const float* U1, const float* r, const float* f,
              const float* c, int nr, int nf, int nc, float* out) {
  const int NS = ModeT<MODE>::nspec;
  const int NC = coef_count(MODE, itype);
  const int nrp = padded_stations(nr);
  // station arrays padded as on the device
  const int ntot = nrp + kUnroll;
  std::vector<double> rr(ntot);
  std::vector<float> hh(ntot), ir(ntot), isr(ntot), dr(ntot);
  std::vector<uint32_t> bw(ntot);
  for (int s = 0; s < ntot; s++) {
    station_entry(r, nr, s, rr.data(), hh.data(), ir.data(), isr.data(), dr.data());
    bw[s] = station_block_word(r, nr, s);
  }
  const Stations S{rr.data(), hh.data(), ir.data(), isr.data(), bw.data(), dr.data()};
  const size_t block = coef_block(nrp, MODE, itype);
  std::vector<float> cf(block, 0.0f);
  float cmin = INFINITY;
  for (int j = 0; j < nc; j++) cmin = fminf(cmin, fabsf(c[j]));
#pragma omp parallel for schedule(dynamic, 1) firstprivate(cf)
  for (int i = 0; i < nf; i++) {
    const float hlim = itype == 1 ? long_segment_limit(f[i], cmin) : 0.0f;
    cf[coef_header(nrp, MODE, itype)] = hlim;
    cf[coef_header(nrp, MODE, itype) + 1] = anchor_limit(f[i], cmin);
    for (int s = 0; s < nr; s++) {
      float gp[2] = {0.f, 0.f}, gc[2] = {0.f, 0.f}, gn[2] = {0.f, 0.f}, rec[12];
      for (int m = 0; m < NS; m++) {
        const float* U = m == 0 ? U0 : U1;
        gc[m] = U[(size_t)s * nf + i];
        if (s > 0) gp[m] = U[(size_t)(s - 1) * nf + i];
        if (s + 1 < nr) gn[m] = U[(size_t)(s + 1) * nf + i];
      }
      const float rp = s > 0 ? r[s - 1] : 0.f, rn = s + 1 < nr ? r[s + 1] : 0.f;
      if (itype == 0) {
        if (NS == 1) station_coef<0, 1>(s, nr, rp, r[s], rn, gp, gc, gn, hlim, rec);
        else station_coef<0, 2>(s, nr, rp, r[s], rn, gp, gc, gn, hlim, rec);
        if (polar_records(MODE, 0)) to_polar(rec);
        if (NS == 1) coef_store<0, 1>(cf.data(), nrp, s, rec);
        else coef_store<0, 2>(cf.data(), nrp, s, rec);
      } else {
        if (NS == 1) station_coef<1, 1>(s, nr, rp, r[s], rn, gp, gc, gn, hlim, rec);
        else station_coef<1, 2>(s, nr, rp, r[s], rn, gp, gc, gn, hlim, rec);
        if (NS == 1) coef_store<1, 1>(cf.data(), nrp, s, rec);
        else coef_store<1, 2>(cf.data(), nrp, s, rec);
      }
    }
    for (int j = 0; j < nc; j++) {
      const double q = (double)f[i] / (double)c[j];
      out[(size_t)j * nf + i] = itype == 0 ? trapz_lane<MODE>(S, cf.data(), nrp, q)
                                           : linseg_lane<MODE>(S, cf.data(), nrp, q);
    }
  }
}

}  // namespace

extern "C" int fjemul_transform(int mode, int itype, const float* U0, const float* U1,
                                const float* r, const float* f, const float* c, int nr, int nf,
                                int nc, float* out) {
  if (itype != 0 && itype != 1) return CCFJ_EINVAL;
  if (nr < 2) {   // no segment: zeros, as the device path
    for (size_t k = 0; k < (size_t)nf * nc; k++) out[k] = 0.f;
    return 0;
  }
  switch (mode) {
    case MODE_J: run_mode<MODE_J>(itype, U0, U1, r, f, c, nr, nf, nc, out); break;
    case MODE_Y: run_mode<MODE_Y>(itype, U0, U1, r, f, c, nr, nf, nc, out); break;
    case MODE_J_MINUS_Y: run_mode<MODE_J_MINUS_Y>(itype, U0, U1, r, f, c, nr, nf, nc, out); break;
    case MODE_ABS_J: run_mode<MODE_ABS_J>(itype, U0, U1, r, f, c, nr, nf, nc, out); break;
    case MODE_ABS_H: run_mode<MODE_ABS_H>(itype, U0, U1, r, f, c, nr, nf, nc, out); break;
    default: return CCFJ_EINVAL;
  }
  return 0;
}

// Scalar access to the special-function layer for unit tests, through the
// same code as the kernels (a lane with k = x at a station at r = 1).
// which: 0 J0, 1 J1, 2 Y0, 3 Y1, 4 G_J, 5 G_Y (x > 0).
extern "C" double fjemul_special(int which, double x) {
  const LaneK L = lane_consts(x / fjm::kTwoPi);
  const float xf = (float)(L.q2pi * 1.0);
  float j0, j1, y0, y1, gj = 0.f, gy = 0.f;
  if (xf >= kLaneFar) {
    // the far stations' forms: modulus-phase for the J-only bases, P and Q
    // for the bases with Y; G from the Struve fits (linseg_station)
    const float z = L.zk, t = z * z, amp = L.ak;
    const float ph = phase_rev(L.q, 1.0);
    const FarUnits wj = far_units<false>(ph, z), wy = far_units<true>(ph, z);
    j0 = amp * wj.u0;
    j1 = amp * wj.u1;
    y0 = amp * wy.v0;
    y1 = amp * wy.v1;
    const float r0 = fit_R0(t), xr1 = xf * fit_R1(t);
    gj = __builtin_fmaf(r0, j1, __builtin_fmaf(-xr1, j0, 1.0f));
    gy = __builtin_fmaf(r0, y1, -xr1 * y0);
  } else {
    double y, Lg;
    const Bess b = fjm::bessel_small<true, true>(x, &y, &Lg);
    j0 = b.j0;
    j1 = b.j1;
    y0 = b.y0;
    y1 = b.y1;
    fjm::g_small(x, y, Lg, true, true, &gj, &gy);
    gy = (float)(gy - fjm::kTwoOverPiD * x);  // unshift
  }
  switch (which) {
    case 0: return j0;
    case 1: return j1;
    case 2: return y0;
    case 3: return y1;
    case 4: return gj;
    case 5: return gy;
  }
  return 0.0;
}
