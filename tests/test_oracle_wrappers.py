"""Pins the oracle's restatement of the reference's Python layer
(oracle/oracle.py win, fj_noise, fj_earthquake, MWFJ; src/ccfj.pyx:148-268)
by facts that do not come from that restatement: window vectors worked out
by hand, scipy.signal's analytic-signal Hilbert transform, the Hankel
identity H0 = J0 + i Y0, linearity, invariance under row permutation and
positive scaling, and the dispersion ridge that Bessel orthogonality puts at
the model velocity.  Each is chosen so that a plausible slip (ramp index,
Hilbert sign, stride before the Hilbert transform, a Re/Im swap, a missing
normalisation) fails one of them.  No GPU."""
import numpy as np
import pytest
from scipy import signal
from scipy.special import j0

import ccfj
from oracle import oracle as O

S1, S2, S3 = np.sin(np.pi / 8), np.sin(np.pi / 4), np.sin(3 * np.pi / 8)
R3 = np.sqrt(3) / 2

# (npts, Fs, T1, T2, taper) -> the window, worked out from src/ccfj.pyx:148-158
HAND_WINDOWS = [
    # regular: n1 = 2, n2 = 8, delta = floor(6 * 0.5 / 2) = 1 -> the ramp is sin(0)
    ((10, 1.0, 2.0, 8.0, 0.5), [0, 0, 0, 1, 1, 1, 1, 0, 0, 0]),
    # taper 0: n1 = 2, n2 = 10, delta = 4: sin(pi/2 i/4), i = 0..3, mirrored
    ((12, 2.0, 1.0, 5.0, 0.0), [0, 0, 0, S1, S2, S3, S3, S2, S1, 0, 0, 0]),
    # taper 1: a box
    ((9, 1.0, 3.0, 7.0, 1.0), [0, 0, 0, 1, 1, 1, 1, 0, 0]),
    # reversed: n1 = 8 > n2 = 2, delta = floor(-1.5) = -2: no ramp, all zero
    ((10, 1.0, 8.0, 2.0, 0.5), [0] * 10),
    # out of range on both sides: n1 = 0, n2 = 10, delta = floor(10 * 0.4 / 2) = 2
    ((10, 1.0, -3.0, 20.0, 0.6), [0, S2, 1, 1, 1, 1, 1, 1, S2, 0]),
    # overlapping ramps (taper -1): n1 = 2, n2 = 5, delta = 3; the second
    # ramp's writes of i = 2 land last
    ((8, 1.0, 2.0, 5.0, -1.0), [0, 0, R3, 0.5, R3, 0, 0, 0]),
    # reversed with a positive delta (taper 3): n1 = 5, n2 = 2, delta = 3;
    # n2 - i - 1 reaches -1, Python's last element
    ((8, 1.0, 5.0, 2.0, 3.0), [0.5, 0, 0, 0, 0, 0, 0.5, R3]),
]


@pytest.mark.parametrize("args,want", HAND_WINDOWS)
def test_win_hand_worked(args, want):
    for fn in (O.win, ccfj.win):
        got = fn(*args)
        assert got.dtype == np.float64 and got.shape == (args[0],)
        assert np.allclose(got, want, rtol=0, atol=1e-15), (fn.__module__, args, got)


def test_win_float32_edges_round_in_float32():
    """Window edges held in float32 (as a float32 winl array hands them to
    win()) make n1, n2 and delta float32 arithmetic: 0.29 s at 100 Hz is
    sample 29 (float64 gives 28), and (329 - 29) * (1 - 0.9) / 2 is 15.0
    (float64 gives 14.999...)."""
    T1, T2 = np.float32(0.29), np.float32(3.29)
    for fn in (O.win, ccfj.win):
        w = fn(500, 100.0, T1, T2, 0.9)
        assert np.all(w[:30] == 0) and w[30] > 0          # n1 = 29, ramp starts at sin(0)
        assert np.all(w[44:314] == 1) and w[43] < 1       # delta = 15
        assert w[314] < 1 and w[328] == 0 and np.all(w[329:] == 0)
        assert abs(w[30] - np.sin(np.pi / 30)) < 1e-7      # float32 ramp
    n1, n2, dl, xidx, xrows = ccfj._window_descriptions(
        np.array([[T1]]), np.array([[T2]]), np.array([0]), 500, 100.0, 0.9)
    assert (n1[0, 0], n2[0, 0], dl[0, 0], xidx[0, 0]) == (29, 329, 15, -1)
    # the same edges in float64: the other sample counts
    n1, n2, dl, _, _ = ccfj._window_descriptions(
        np.array([[float(T1)]]), np.array([[float(T2)]]), np.array([0]), 500, 100.0, 0.9)
    assert (n1[0, 0], n2[0, 0], dl[0, 0]) == (28, 328, 14)


def test_window_descriptions_match_win_elementwise():
    """The integer descriptions MWFJ ships to the device rebuild, for float32
    and float64 edges, exactly the window the element-wise restatement
    builds (regular windows from (n1, n2, delta), degenerate ones as rows)."""
    rng = np.random.default_rng(8)
    npts, Fs = 300, 10.0
    for dt in (np.float32, np.float64):
        winl = (rng.uniform(-5, 30, (6, 40))).astype(dt)
        winr = (winl + rng.uniform(-4, 25, (6, 40))).astype(dt)
        for taper in (0.9, 0.8, 0.0):
            n1, n2, dl, xidx, xrows = ccfj._window_descriptions(winl, winr, np.arange(40), npts,
                                                                Fs, taper)
            for w in range(6):
                for s in range(40):
                    want = O.win(npts, Fs, winl[w, s], winr[w, s], taper)
                    if xidx[w, s] >= 0:
                        got = xrows[xidx[w, s]]
                    else:
                        got = np.zeros(npts)
                        a, b, d = n1[w, s], n2[w, s], dl[w, s]
                        got[a:b] = 1
                        i = np.arange(d)
                        got[a + i] = np.sin(np.pi / 2 * i / d)
                        got[b - 1 - i] = np.sin(np.pi / 2 * i / d)
                    # the device evaluates ramps in double; float32 edges
                    # give the reference float32 ramp values
                    tol = 1e-7 if dt == np.float32 else 1e-15
                    assert np.allclose(got, want, rtol=0, atol=tol), (dt, taper, w, s)


# ------------------------------------------------------------------ fj_noise
@pytest.fixture(scope="module")
def small():
    rng = np.random.default_rng(21)
    nr, ncol = 40, 61
    r = rng.uniform(2e3, 80e3, nr)
    f = np.linspace(0.02, 0.6, ncol)
    c0 = 2500.0 + 1200.0 * np.exp(-f / 0.15)
    uf = j0(2 * np.pi * f[None, :] * r[:, None] / c0[None, :]) + rng.normal(0, 0.2, (nr, ncol))
    return dict(uf=uf.astype(np.float32), r=r, f=f, c=np.linspace(1500.0, 5000.0, 90), c0=c0)


@pytest.mark.parametrize("itype", [0, 1])
def test_fj_noise_hilbert_of_full_rows_then_stride(small, itype):
    """func 1 = N(FJ[U] - FY[H U]) with H the Hilbert transform of the full
    row along frequency, taken before the column stride; here H is
    -Im(scipy.signal.hilbert) (fftpack's sign: hilbert(cos) = -sin)."""
    d = small
    idx = np.argsort(d["r"])
    uf, r = d["uf"][idx], d["r"][idx]
    fs = d["f"][::3]
    H = (-np.imag(signal.hilbert(uf.astype(np.float64), axis=1))).astype(np.float32)
    J = O.oracle64(uf[:, ::3], r, fs, d["c"], itype, 0)
    right = J - O.oracle64(H[:, ::3], r, fs, d["c"], itype, 1)
    right = right / np.max(np.abs(right), axis=0)
    got = O.fj_noise(d["uf"], d["r"], d["c"], fs, fstride=3, itype=itype, func=1)
    assert np.max(np.abs(got - right)) < 1e-6
    # the other order (stride, then transform) is a different answer
    Hs = (-np.imag(signal.hilbert(uf[:, ::3].astype(np.float64), axis=1))).astype(np.float32)
    wrong = J - O.oracle64(Hs, r, fs, d["c"], itype, 1)
    wrong = wrong / np.max(np.abs(wrong), axis=0)
    assert np.max(np.abs(got - wrong)) > 1e-2


@pytest.mark.parametrize("func", [0, 1])
def test_fj_noise_properties(small, func):
    """Row permutation (the sort), positive scaling (the normalisation),
    linearity before the normalisation, and column maxima of one."""
    d = small
    base = O.fj_noise(d["uf"], d["r"], d["c"], d["f"], itype=1, func=func)
    perm = np.random.default_rng(2).permutation(len(d["r"]))
    assert np.array_equal(O.fj_noise(d["uf"][perm], d["r"][perm], d["c"], d["f"], itype=1,
                                     func=func), base)
    scaled = O.fj_noise(d["uf"] * np.float32(4.0), d["r"], d["c"], d["f"], itype=1, func=func)
    assert np.max(np.abs(scaled - base)) < 1e-12
    assert np.allclose(np.max(np.abs(base), axis=0), 1.0, rtol=0, atol=1e-15)
    V = np.roll(d["uf"], 5, axis=1)
    a = O.fj_noise(d["uf"] + 2 * V, d["r"], d["c"], d["f"], itype=1, func=func, normalize=False)
    b = (O.fj_noise(d["uf"], d["r"], d["c"], d["f"], itype=1, func=func, normalize=False)
         + 2 * O.fj_noise(V, d["r"], d["c"], d["f"], itype=1, func=func, normalize=False))
    assert np.max(np.abs(a - b)) < 1e-6 * np.max(np.abs(b))
    # normalisation really divides by the column max: unnormalised / max == base
    raw = O.fj_noise(d["uf"], d["r"], d["c"], d["f"], itype=1, func=func, normalize=False)
    assert np.max(np.abs(raw / np.max(np.abs(raw), axis=0) - base)) < 1e-15


@pytest.mark.parametrize("itype", [0, 1])
@pytest.mark.parametrize("func", [0, 1])
def test_fj_noise_ridge_at_model_velocity(itype, func):
    """U = J0(2 pi f r / c0(f)) on a dense array: by the orthogonality of
    J0 over r, the column maximum of both bases' spectra sits at c0(f)."""
    rng = np.random.default_rng(4)
    r = rng.uniform(5e3, 250e3, 400)
    f = np.linspace(0.06, 0.3, 13)
    c0 = 3000.0 + 1500.0 * np.exp(-f / 0.08)
    uf = j0(2 * np.pi * f[None, :] * r[:, None] / c0[None, :]).astype(np.float32)
    c = np.linspace(2000.0, 5000.0, 301)
    out = O.fj_noise(uf, r, c, f, itype=itype, func=func)
    pick = c[np.argmax(out, axis=0)]
    assert np.max(np.abs(pick - c0)) <= 3 * (c[1] - c[0]), (pick, c0)


# ------------------------------------------------------------- fj_earthquake
def _records(seed=6, nr=30, npts=256):
    rng = np.random.default_rng(seed)
    r = rng.uniform(1e3, 40e3, nr)
    return rng.standard_normal((nr, npts)).astype(np.float32), r


@pytest.mark.parametrize("itype", [0, 1])
def test_fj_earthquake_by_hand(itype):
    """func 0: |FJ[Re X] + i FJ[Im X]|; func 1: |sum X H0(k r) r| with
    H0 = J0 + i Y0 applied to the complex spectrum X = rfft(u), computed
    here in complex arithmetic from the single-basis oracle."""
    u, r = _records()
    f = np.arange(1, 41) * 0.02
    c = np.linspace(1500.0, 4000.0, 50)
    idx = np.argsort(r)
    X = np.fft.rfft(u[idx])[:, :len(f)]
    Re = np.ascontiguousarray(X.real, np.float32)
    Im = np.ascontiguousarray(X.imag, np.float32)
    rs = r[idx]
    J = O.oracle64(Re, rs, f, c, itype, 0) + 1j * O.oracle64(Im, rs, f, c, itype, 0)
    Y = O.oracle64(Re, rs, f, c, itype, 1) + 1j * O.oracle64(Im, rs, f, c, itype, 1)
    for func, want in ((0, np.abs(J)), (1, np.abs(J + 1j * Y))):
        want = want / np.max(want, axis=0)
        got = O.fj_earthquake(u, r, c, f, itype=itype, func=func)
        assert np.max(np.abs(got - want)) < 1e-12, func


def test_mwfj_windows():
    """A window covering the whole trace is fj_earthquake itself; a window
    [n1, n2) with taper 1 is fj_earthquake of the traces zeroed outside it."""
    u, r = _records(nr=12, npts=200)
    Fs = 10.0
    f = np.arange(1, 30) * Fs / 200
    c = np.linspace(1500.0, 4000.0, 40)
    winl = np.stack([np.full(12, -1.0), np.full(12, 3.0)])
    winr = np.stack([np.full(12, 100.0), np.full(12, 12.0)])
    out = O.MWFJ(u, r, c, f, Fs, 2, winl, winr, taper=1.0, itype=0)
    assert out.dtype == np.float64 and out.shape == (2, len(c), len(f))
    assert np.array_equal(out[0], O.fj_earthquake(u.astype(np.float64), r, c, f, itype=0))
    cut = np.zeros_like(u, dtype=np.float64)
    cut[:, 30:120] = u[:, 30:120]
    assert np.array_equal(out[1], O.fj_earthquake(cut, r, c, f, itype=0))
