"""The F-J kernel arithmetic, compiled for the host (libccfj_emul.so from
fj_lane.h / fj_math.h -- the same source the gfx950 kernels are built
from), checked against the FP64 oracle and mpmath/scipy without a GPU.

The host build uses portable sin/cos/rsq/log in place of the hardware
instructions; everything else (coefficients, phase reduction, the
small/large split, the segment rule, accumulation) is the kernel's code.
"""
import ctypes
import os

import mpmath as mp
import numpy as np
import pytest
from scipy import special as sp

from conftest import GOLDEN_CASES, ROOT, golden_case
from oracle import oracle as O

TOL = 1e-5
MODE_J, MODE_Y, MODE_JMY, MODE_ABSJ, MODE_ABSH = 0, 1, 2, 3, 4


@pytest.fixture(scope="module")
def emul():
    path = os.path.join(ROOT, "cc-fjpy_amd", "ccfj", "libccfj_emul.so")
    if not os.path.exists(path):
        import __graft_entry__
        __graft_entry__.build()
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    lib.fjemul_transform.argtypes = [ctypes.c_int, ctypes.c_int] + [P] * 5 + [ctypes.c_int] * 3 + [P]
    lib.fjemul_special.argtypes = [ctypes.c_int, ctypes.c_double]
    lib.fjemul_special.restype = ctypes.c_double
    return lib


def run(lib, mode, itype, U0, r, f, c, U1=None):
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    U0, r, f, c = f32(U0), f32(r), f32(f), f32(c)
    U1 = U0 if U1 is None else f32(U1)
    out = np.zeros((len(c), len(f)), np.float32)
    rc = lib.fjemul_transform(mode, itype, U0.ctypes.data, U1.ctypes.data, r.ctypes.data,
                              f.ctypes.data, c.ctypes.data, len(r), len(f), len(c),
                              out.ctypes.data)
    assert rc == 0
    return out


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("ybasis", [0, 1])
@pytest.mark.parametrize("itype", [0, 1])
def test_golden_vs_oracle(emul, golden, case, ybasis, itype):
    U, r, f, c = golden_case(golden, case)
    got = run(emul, MODE_Y if ybasis else MODE_J, itype, U, r, f, c)
    emax, erms, mism = O.column_error(got, O.oracle64(U, r, f, c, itype, ybasis))
    assert mism == 0
    assert emax <= TOL


@pytest.mark.parametrize("itype", [0, 1])
def test_fused_modes(emul, golden, itype):
    U, r, f, c = golden_case(golden, "branches")
    V = np.flipud(U).copy() * 0.3
    J0, J1 = run(emul, MODE_J, itype, U, r, f, c), run(emul, MODE_J, itype, V, r, f, c)
    Y0, Y1 = run(emul, MODE_Y, itype, U, r, f, c), run(emul, MODE_Y, itype, V, r, f, c)
    J0, J1, Y0, Y1 = (a.astype(np.float64) for a in (J0, J1, Y0, Y1))
    for mode, want in ((MODE_JMY, J0 - Y1), (MODE_ABSJ, np.hypot(J0, J1)),
                       (MODE_ABSH, np.hypot(J0 - Y1, J1 + Y0))):
        got = run(emul, mode, itype, U, r, f, c, U1=V)
        assert O.column_error(got, want)[0] < 1e-6, mode


@pytest.mark.parametrize("itype", [0, 1])
def test_anchored_block_phases(emul, itype):
    """One-spectrum modes take a block's phases from one double-precision
    reduction at its first station plus q (r_s - r_a) in float when the
    block's spread of r is within the frequency's anchor limit (1 / max q,
    250 m at 2 Hz over c >= 500 m/s here): error at most 1.5e-7
    revolutions per phase.  Spacings of 80 and 90 m put the blocks of 4
    (spans 240 and 270 m) just inside and just outside the limit at 2 Hz
    (both inside at 1 Hz); against the FP64 oracle the results stay within
    3e-6.  At 800 m the blocks are 10 times the limit and must take the
    per-station reduction: a limit that let them anchor would add errors of
    several 1e-6 (|q dr| up to 10 revolutions in float), above the 1e-6
    asserted there."""
    rng = np.random.default_rng(3)
    f = np.array([1.0, 2.0], np.float32)
    c = np.linspace(500, 3000, 96).astype(np.float32)
    for d, bound in ((80.0, 3e-6), (90.0, 3e-6), (800.0, 1e-6)):
        r = (20e3 + d * np.arange(400)).astype(np.float32)
        U = (rng.normal(size=(400, 2)) + 1.0).astype(np.float32)
        for mode, ybasis in ((MODE_J, 0), (MODE_Y, 1)):
            got = run(emul, mode, itype, U, r, f, c)
            emax, _, mism = O.column_error(got, O.oracle64(U, r, f, c, itype, ybasis))
            assert mism == 0
            assert emax <= bound, (d, mode, emax)


def test_special_functions(emul):
    """J0, J1, Y0, Y1 and G(x) = int_0^x t Z1 dt against scipy/mpmath:
    absolute error (G: relative to max(1, sqrt x)) across both branches."""
    xs = np.concatenate([np.geomspace(1e-3, 7.999, 300), np.geomspace(8.0, 2e4, 600)])
    G = lambda Z0, Z1, x: np.pi * x / 2 * (Z1(x) * sp.struve(0, x) - Z0(x) * sp.struve(1, x))
    ref = [sp.j0(xs), sp.j1(xs), sp.y0(xs), sp.y1(xs),
           G(sp.j0, sp.j1, xs), G(sp.y0, sp.y1, xs)]
    for w in range(6):
        got = np.array([emul.fjemul_special(w, x) for x in xs])
        scale = np.maximum(1.0, np.sqrt(xs)) if w >= 4 else np.maximum(1.0, np.abs(ref[w]))
        ok = np.isfinite(ref[w]) & (xs < 3e3 if w >= 4 else True)   # scipy struve loses digits beyond
        err = np.abs(got - ref[w])[ok] / scale[ok]
        assert err.max() < 3e-7, (w, err.max(), xs[ok][np.argmax(err)])
    mp.mp.dps = 30
    for x in (3e3, 7.5e3, 1.1e4):
        gj = float(mp.quad(lambda t: t * mp.besselj(1, t), mp.linspace(0, x, 400)))
        assert abs(emul.fjemul_special(4, x) - gj) < 3e-7 * np.sqrt(x)


@pytest.mark.parametrize("itype", [0, 1])
def test_dense_column_vs_oracle(emul, itype):
    """One low and one mid frequency of the 19,900-pair dense array, all
    3000 velocities: the accumulation over ~2e4 stations with mostly
    Euler-Maclaurin segments stays within the bar."""
    from ccfj import synth
    fidx = [7, 1500]
    d = synth.noise_dense(real_only=True, fidx=fidx)
    idx = np.argsort(d["r"])
    U, r = d["uf"][idx], d["r"][idx].astype(np.float32)
    c = d["c"].astype(np.float32)
    f = d["f"].astype(np.float32)
    got = run(emul, MODE_J, itype, U, r, f, c)
    emax, erms, mism = O.column_error(got, O.oracle64(U, r, f, c, itype, 0))
    assert mism == 0
    assert emax <= TOL


@pytest.mark.parametrize("mode", [MODE_J, MODE_Y, MODE_JMY])
def test_long_and_short_segments(emul, mode):
    """Linear segments on an array that mixes clusters of short gaps with
    gaps of tens of kilometres: over the frequency grid, k_max h runs from
    far below to far above the long-segment limit (2), so every kind of
    segment occurs -- in the per-station records, recorded but taking the G
    difference on some lanes, and long (summed per lane) -- also between
    neighbours of different kinds."""
    rng = np.random.default_rng(11)
    parts = [np.sort(rng.uniform(a, a + w, n)) for a, w, n in
             ((2e3, 3e3, 40), (30e3, 1e3, 25), (75e3, 40e3, 12), (160e3, 2e3, 30))]
    r = np.concatenate(parts).astype(np.float32)
    nr = len(r)
    f = np.linspace(0.0, 1.5, 24).astype(np.float32)
    c = np.linspace(300.0, 5000.0, 40).astype(np.float32)
    U0 = rng.normal(size=(nr, len(f))).astype(np.float32)
    U1 = rng.normal(size=(nr, len(f))).astype(np.float32)
    got = run(emul, mode, 1, U0, r, f, c, U1=U1)
    if mode == MODE_JMY:
        want = O.oracle64(U0, r, f, c, 1, 0) - O.oracle64(U1, r, f, c, 1, 1)
    else:
        want = O.oracle64(U0, r, f, c, 1, 1 if mode == MODE_Y else 0)
    emax, _, mism = O.column_error(got[:, 1:], want[:, 1:])   # f = 0: NaN in both
    assert mism == 0
    assert emax <= TOL


def test_unsorted_stations_same_math(emul, golden):
    """Unsorted r (the C level does not sort) still gives the integral the
    reference's formulas define (negative gaps are clamped to 0.1 m)."""
    U, r, f, c = golden_case(golden, "cfg1s")
    perm = np.random.default_rng(2).permutation(len(r))[:40]
    for itype in (0, 1):
        got = run(emul, MODE_J, itype, U[perm], r[perm], f, c)
        emax, _, mism = O.column_error(got, O.oracle64(U[perm], r[perm], f, c, itype, 0))
        assert mism == 0 and emax <= 1e-4
