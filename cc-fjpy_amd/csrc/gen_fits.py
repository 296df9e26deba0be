"""Generate fj_fits.h: polynomial fits of the Bessel/Struve building blocks.

    python cc-fjpy_amd/csrc/gen_fits.py      # rewrites fj_fits.h (about 1 min)

Every fit is a least-squares Chebyshev fit (Chebyshev nodes, mpmath ground
truth at 40 digits) converted to monomial form and evaluated by Horner's rule
with fused multiply-adds.  Nothing is taken from other libraries' tables.

Large argument, x >= 8, variable t = (8/x)^2 in (0,1], evaluated in float:
  Hankel asymptotic form (Abramowitz & Stegun 9.2.5-9.2.6):
    J_n(x) = sqrt(2/(pi x)) (P_n cos th_n - Q_n sin th_n)
    Y_n(x) = sqrt(2/(pi x)) (P_n sin th_n + Q_n cos th_n),  th_n = x - (2n+1) pi/4
  fit_P0(t) = P_0(x), fit_q0(t) = Q_0(x) x/8, fit_P1, fit_q1 likewise.
  Struve minus Neumann (A&S 12.1.29-12.1.31), smooth and non-oscillating:
    fit_R0(t) = (pi/2) x (H_0(x) - Y_0(x)),  fit_R1(t) = (pi/2) (H_1(x) - Y_1(x))
Far argument, x >= 20 (t <= 0.16): the same six functions at lower degree,
fitf_P0 ... fitf_R1 (degree 2 for P and Q, 3 for R; relative error below
1e-8 on that range).
Far argument, modulus-phase form (one transcendental per function instead
of a sine and a cosine):
    J_n(x) = sqrt(2/(pi x)) M_n cos(th_n + phi_n),  Y_n = ... sin(th_n + phi_n),
    M_n = sqrt(P_n^2 + Q_n^2),  phi_n = atan2(Q_n, P_n)
  fitf_M0(t) = M_0,  fitf_ph0(t) = phi_0 / (2 pi) * x/8  (degree 2 and 1),
  fitf_M1, fitf_ph1 likewise for n = 1.
From x = 8 on (t in (0, 1]), for the F-J kernels' far stations: fitm_M0 ...
fitm_ph1 and fitm_P0 ... fitm_q1, all degree 2: errors up to 8e-8 in the
moduli and P, 2e-8 in Q/z and 4e-9 revolutions x 8/x in the phases, below
the float rounding of the terms they enter (degree 3 is about 2e-9).
Small argument, x < 8, variable y = (x/8)^2 in [0,1), evaluated in double
(the monomial coefficients alternate in sign up to ~1e2, too much
cancellation for float):
  J_0 = fit_J0s(y);  J_1 = x fit_J1s(y)
  Y_0 = (2/pi) ln(x) J_0 + fit_Y0s(y)
  Y_1 = (2/pi) ln(x) J_1 - 2/(pi x) + x fit_Y1s(y)
  G_J(x) = int_0^x t J_1(t) dt = x^3 fit_GJs(y)
  G_Y(x) = int_0^x t Y_1(t) dt = (2/pi) ln(x) G_J(x) + x fit_GYs(y)
"""
import os

import mpmath as mp
import struct

import numpy as np

mp.mp.dps = 40
P = np.polynomial
HERE = os.path.dirname(os.path.abspath(__file__))


def cheb_nodes(n, a, b):
    k = np.arange(n)
    return (a + b) / 2 + (b - a) / 2 * np.cos(np.pi * (k + 0.5) / n)


def large_truth(x):
    x = mp.mpf(x)
    J0, Y0 = mp.besselj(0, x), mp.bessely(0, x)
    J1, Y1 = mp.besselj(1, x), mp.bessely(1, x)
    th0, th1 = x - mp.pi / 4, x - 3 * mp.pi / 4
    s = mp.sqrt(mp.pi * x / 2)
    z = 8 / x
    P0 = s * (J0 * mp.cos(th0) + Y0 * mp.sin(th0))
    Q0 = s * (Y0 * mp.cos(th0) - J0 * mp.sin(th0))
    P1 = s * (J1 * mp.cos(th1) + Y1 * mp.sin(th1))
    Q1 = s * (Y1 * mp.cos(th1) - J1 * mp.sin(th1))
    S0 = mp.struveh(0, x) - Y0
    S1 = mp.struveh(1, x) - Y1
    return [P0, Q0 / z, P1, Q1 / z, mp.pi / 2 * x * S0, mp.pi / 2 * S1]


def far_mp_truth(x):
    v = large_truth(x)
    z = 8 / mp.mpf(x)
    P0, Q0 = v[0], v[1] * z
    P1, Q1 = v[2], v[3] * z
    return [mp.sqrt(P0 ** 2 + Q0 ** 2), mp.atan2(Q0, P0) / (2 * mp.pi) / z,
            mp.sqrt(P1 ** 2 + Q1 ** 2), mp.atan2(Q1, P1) / (2 * mp.pi) / z]


def small_truth(x):
    x = mp.mpf(x)
    L = 2 / mp.pi * mp.log(x)
    J0, J1 = mp.besselj(0, x), mp.besselj(1, x)
    Y0, Y1 = mp.bessely(0, x), mp.bessely(1, x)
    H0, H1 = mp.struveh(0, x), mp.struveh(1, x)
    gj = mp.pi * x / 2 * (J1 * H0 - J0 * H1)
    gy = mp.pi * x / 2 * (Y1 * H0 - Y0 * H1)
    return [J0, J1 / x, Y0 - L * J0, (Y1 - L * J1 + 2 / (mp.pi * x)) / x,
            gj / x ** 3, (gy - L * gj) / x]


LARGE = [("fit_P0", 4), ("fit_q0", 4), ("fit_P1", 4), ("fit_q1", 4),
         ("fit_R0", 7), ("fit_R1", 6)]
FAR_X = 20.0
FAR = [("fitf_P0", 2), ("fitf_q0", 2), ("fitf_P1", 2), ("fitf_q1", 2),
       ("fitf_R0", 3), ("fitf_R1", 3)]
FAR_MP = [("fitf_M0", 2), ("fitf_ph0", 1), ("fitf_M1", 2), ("fitf_ph1", 1)]
# the F-J lane's far forms from x = 8 on (t in (0, 1]): modulus-phase and P, Q
MID_MP = [("fitm_M0", 2), ("fitm_ph0", 2), ("fitm_M1", 2), ("fitm_ph1", 2)]
MID_PQ = [("fitm_P0", 2), ("fitm_q0", 2), ("fitm_P1", 2), ("fitm_q1", 2)]
SMALL = [("fit_J0s", 11), ("fit_J1s", 10), ("fit_Y0s", 11), ("fit_Y1s", 10),
         ("fit_GJs", 10), ("fit_GYs", 11)]


def fit_all(names, var, truth_fn, npts=160, top=1.0):
    v = cheb_nodes(npts, 0.0, top)
    v = v[v > 1e-7 * top]
    xs = 8 / np.sqrt(v) if var == "t" else 8 * np.sqrt(v)
    vals = np.array([[float(q) for q in truth_fn(x)] for x in xs])
    # independent check points
    rng = np.random.default_rng(1)
    vc = np.sort(rng.uniform(1e-6 * top, top, 300))
    xc = 8 / np.sqrt(vc) if var == "t" else 8 * np.sqrt(vc)
    valc = np.array([[float(q) for q in truth_fn(x)] for x in xc])
    res = []
    for j, (name, deg) in enumerate(names):
        cheb = P.Chebyshev.fit(v, vals[:, j], deg, domain=[0, top])
        coef = cheb.convert(kind=P.Polynomial, domain=[0, top], window=[0, top]).coef
        err = float(np.max(np.abs(P.polynomial.polyval(vc, coef) - valc[:, j])))
        res.append((name, coef, err))
    return res


def emit(fname, coefs, ctype, err, var, top=1.0):
    fma = "__builtin_fmaf" if ctype == "float" else "__builtin_fma"
    suf = "f" if ctype == "float" else ""
    head = f"/* max abs error of the fit on (0,{top:g}]: {err:.2e} */"
    if ctype == "double":
        # three forms of one polynomial, chosen per kernel by TAB: literal
        # coefficients (FIT_LITERAL), a table (FIT_TABLE; on the device a
        # __constant__ array whose loads the compiler hoists into scalar
        # registers at kernel entry), or scalar moves at the point of use
        # (FIT_SCALAR; FJ_SCONST, which the compiler cannot hoist out of the
        # branch that evaluates the polynomial).  Which is faster depends on
        # the kernel's register pressure.
        vals = ", ".join(f"{float(c)!r}" for c in coefs[::-1])

        def sconst(c):
            bits = struct.unpack("<Q", struct.pack("<d", float(c)))[0]
            return f"FJ_SCONST({float(c)!r}, 0x{bits & 0xffffffff:08x}u, 0x{bits >> 32:08x}u)"

        lines = [head,
                 f"FJ_TABLE(k_{fname}, {len(coefs)}, {{{vals}}})",
                 "template <int TAB = FIT_LITERAL>",
                 f"FJ_HD inline double {fname}(double {var}) {{",
                 "  if (TAB == FIT_TABLE) {",
                 f"    const double* c = FJ_TABLE_PTR(k_{fname});",
                 "    double p = c[0];"]
        for k in range(1, len(coefs)):
            lines.append(f"    p = {fma}(p, {var}, c[{k}]);")
        lines += ["    return p;", "  }", "  if (TAB == FIT_SCALAR) {",
                  f"    double p = {sconst(coefs[-1])};"]
        for c in coefs[-2::-1]:
            lines.append(f"    p = {fma}(p, {var}, {sconst(c)});")
        lines += ["    return p;", "  }", f"  double p = {float(coefs[-1])!r};"]
        for c in coefs[-2::-1]:
            lines.append(f"  p = {fma}(p, {var}, {float(c)!r});")
        lines.append("  return p;")
        lines.append("}")
        return "\n".join(lines)
    lines = [head,
             f"FJ_HD inline {ctype} {fname}({ctype} {var}) {{",
             f"  {ctype} p = {float(coefs[-1])!r}{suf};"]
    for c in coefs[-2::-1]:
        lines.append(f"  p = {fma}(p, {var}, {float(c)!r}{suf});")
    lines.append("  return p;")
    lines.append("}")
    return "\n".join(lines)


def main():
    large = fit_all(LARGE, "t", large_truth)
    far_top = (8.0 / FAR_X) ** 2
    far = fit_all(FAR, "t", large_truth, npts=80, top=far_top)
    far_mp = fit_all(FAR_MP, "t", far_mp_truth, npts=80, top=far_top)
    small = fit_all(SMALL, "y", small_truth)
    mid_mp = fit_all(MID_MP, "t", far_mp_truth, npts=120)
    mid_pq = fit_all(MID_PQ, "t", large_truth, npts=120)
    out = ["/* Generated by gen_fits.py -- do not edit by hand. */",
           "#pragma once",
           "#ifndef FJ_HD",
           "#define FJ_HD",
           "#endif",
           "/* double tables: __constant__ on the device, plain on the host */",
           "enum { FIT_LITERAL = 0, FIT_TABLE = 1, FIT_SCALAR = 2 };",
           "/* FJ_SCONST(v, lo, hi): the double v (bits hi:lo); on the device two",
           "   s_mov_b32 into a scalar register pair where the value is used (a",
           "   volatile asm, so it is neither hoisted nor kept live across a loop) */",
           "#if defined(__HIP_DEVICE_COMPILE__)",
           "template <unsigned LO, unsigned HI>",
           "__device__ inline double fj_sconst() {",
           "  unsigned lo, hi;",
           "  asm volatile(\"s_mov_b32 %0, %2\\n\\ts_mov_b32 %1, %3\" : \"=s\"(lo), \"=s\"(hi) : \"n\"(LO), \"n\"(HI));",
           "  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);",
           "}",
           "#define FJ_SCONST(v, lo, hi) fj_sconst<lo, hi>()",
           "#else",
           "#define FJ_SCONST(v, lo, hi) (v)",
           "#endif",
           "#if defined(__HIPCC__)",
           "#define FJ_TABLE(name, n, ...) static __constant__ double name##_d[n] = __VA_ARGS__; "
           "static const double name##_h[n] = __VA_ARGS__;",
           "#if defined(__HIP_DEVICE_COMPILE__)",
           "#define FJ_TABLE_PTR(name) name##_d",
           "#else",
           "#define FJ_TABLE_PTR(name) name##_h",
           "#endif",
           "#else",
           "#define FJ_TABLE(name, n, ...) static const double name##_h[n] = __VA_ARGS__;",
           "#define FJ_TABLE_PTR(name) name##_h",
           "#endif",
           "",
           "/* large argument x >= 8, t = (8/x)^2, float */"]
    for name, coef, err in large:
        out.append(emit(name, coef, "float", err, "t"))
    out.append("")
    out.append(f"/* far argument x >= {FAR_X:g}, t = (8/x)^2 <= {far_top:g}, float */")
    out.append(f"constexpr float kFarX = {FAR_X!r}f;")
    for name, coef, err in far:
        out.append(emit(name, coef, "float", err, "t", far_top))
    out.append("")
    out.append(f"/* far argument, modulus and phase of the order-0 and order-1 functions:")
    out.append(f"   Mn(t), and the phase shift in revolutions over 8/x */")
    for name, coef, err in far_mp:
        out.append(emit(name, coef, "float", err, "t", far_top))
    out.append("")
    out.append("/* x >= 8 (t in (0, 1]), the F-J lane's far forms: modulus and phase, and P, Q */")
    for name, coef, err in mid_mp + mid_pq:
        out.append(emit(name, coef, "float", err, "t"))
    out.append("")
    out.append("/* small argument x < 8, y = (x/8)^2, double */")
    for name, coef, err in small:
        out.append(emit(name, coef, "double", err, "y"))
    path = os.path.join(HERE, "fj_fits.h")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
    for name, coef, err in large + far + far_mp + mid_mp + mid_pq + small:
        print(f"{name:10s} deg {len(coef) - 1:2d} err {err:.2e}")


if __name__ == "__main__":
    main()
