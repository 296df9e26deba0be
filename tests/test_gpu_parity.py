"""GPU parity of the HIP F-J path (libccfj.so through the C ABI).

Bars (written into each test):
  * vs the FP64 oracle: max over entries of |X - X*| / max_c |X*[:, f]|
    <= 1e-5 (the accuracy target), NaN patterns identical;
  * vs the FJcpu restatement (bit-exact with the reference): within
    2 * (FJcpu's own distance to the oracle) + 1e-5 -- FJcpu's float
    closed form is far from the true integral for itype 1, so it cannot be
    held to 1e-5 itself (see test_oracle.py);
  * wrapper semantics (sort, cast, stride, Hilbert, combine, normalise)
    against the restatement of the reference's Python layer.
"""
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN_CASES, ROOT, golden_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def native(ctx, mode, itype, U0, r, f, c, U1=None, normalize=False):
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    return ctx.transform(mode, itype, normalize, f32(U0), None if U1 is None else f32(U1),
                         f32(r), f32(f), f32(c))


# ---------------------------------------------------------------- golden cases
@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("ybasis", [0, 1])
@pytest.mark.parametrize("itype", [0, 1])
def test_golden_vs_oracle_and_fjcpu(ccfj, ctx, golden, case, ybasis, itype):
    U, r, f, c = golden_case(golden, case)
    got = native(ctx, ccfj.MODE_Y if ybasis else ccfj.MODE_J, itype, U, r, f, c)
    ref64 = O.oracle64(U, r, f, c, itype, ybasis)
    emax, erms, mism = O.column_error(got, ref64)
    print(f"{case} basis={'Y' if ybasis else 'J'} itype={itype}: emax={emax:.2e} erms={erms:.2e}")
    assert mism == 0
    assert emax <= TOL
    fjcpu = golden[f"{case}/{'fh' if ybasis else 'fj'}{itype}"]
    e_ref, _, _ = O.column_error(fjcpu, ref64)
    e_cmp, _, mism2 = O.column_error(got, fjcpu)
    assert mism2 == 0
    assert e_cmp <= 2 * e_ref + TOL


def test_golden_vs_host_build(ccfj, ctx, golden):
    """The GPU kernels and the host build of the same code agree up to the
    hardware sin/cos/rsq/log instructions (~1e-7)."""
    import ctypes
    lib = ctypes.CDLL(f"{ROOT}/cc-fjpy_amd/ccfj/libccfj_emul.so")
    P = ctypes.c_void_p
    lib.fjemul_transform.argtypes = [ctypes.c_int, ctypes.c_int] + [P] * 5 + [ctypes.c_int] * 3 + [P]
    for case in GOLDEN_CASES:
        U, r, f, c = golden_case(golden, case)
        for mode in (ccfj.MODE_J, ccfj.MODE_Y):
            for itype in (0, 1):
                host = np.zeros((len(c), len(f)), np.float32)
                lib.fjemul_transform(mode, itype, U.ctypes.data, U.ctypes.data, r.ctypes.data,
                                     f.ctypes.data, c.ctypes.data, len(r), len(f), len(c),
                                     host.ctypes.data)
                got = native(ctx, mode, itype, U, r, f, c)
                emax, _, mism = O.column_error(got, host)
                assert mism == 0 and emax < 2e-6, (case, mode, itype, emax)


# ------------------------------------------------------------- fused modes
@pytest.mark.parametrize("itype", [0, 1])
def test_fused_modes_match_single_basis(ccfj, ctx, golden, itype):
    U, r, f, c = golden_case(golden, "cfg1s")
    V = np.roll(U, 3, axis=0) * 0.7
    J0, J1 = native(ctx, ccfj.MODE_J, itype, U, r, f, c), native(ctx, ccfj.MODE_J, itype, V, r, f, c)
    Y0, Y1 = native(ctx, ccfj.MODE_Y, itype, U, r, f, c), native(ctx, ccfj.MODE_Y, itype, V, r, f, c)
    J0, J1, Y0, Y1 = (a.astype(np.float64) for a in (J0, J1, Y0, Y1))
    cases = {
        ccfj.MODE_J_MINUS_Y: J0 - Y1,
        ccfj.MODE_ABS_J: np.hypot(J0, J1),
        ccfj.MODE_ABS_H: np.hypot(J0 - Y1, J1 + Y0),
    }
    for mode, want in cases.items():
        got = native(ctx, mode, itype, U, r, f, c, U1=V)
        emax, _, mism = O.column_error(got, want)
        assert mism == 0 and emax < 2e-6, (mode, emax)


# ------------------------------------------------------- single-array workload
@pytest.mark.parametrize("mode", ["J", "Y", "J_MINUS_Y"])
def test_long_and_short_segments(ccfj, ctx, mode):
    """Linear segments on an array of clusters of short gaps separated by
    gaps of tens of kilometres, 256 velocities per frequency: the kernel
    meets smooth far blocks, far blocks with long segments or G
    differences, and near blocks, and every kind of segment, against the
    FP64 oracle."""
    rng = np.random.default_rng(11)
    parts = [np.sort(rng.uniform(a, a + w, n)) for a, w, n in
             ((2e3, 3e3, 40), (30e3, 1e3, 25), (75e3, 40e3, 12), (160e3, 2e3, 30))]
    r = np.concatenate(parts).astype(np.float32)
    f = np.linspace(0.0, 1.5, 24).astype(np.float32)
    c = np.linspace(300.0, 5000.0, 256).astype(np.float32)
    U0 = rng.normal(size=(len(r), len(f))).astype(np.float32)
    U1 = rng.normal(size=(len(r), len(f))).astype(np.float32)
    m = getattr(ccfj, "MODE_" + mode)
    got = native(ctx, m, 1, U0, r, f, c, U1=U1)
    if mode == "J_MINUS_Y":
        want = O.oracle64(U0, r, f, c, 1, 0) - O.oracle64(U1, r, f, c, 1, 1)
    else:
        want = O.oracle64(U0, r, f, c, 1, 1 if mode == "Y" else 0)
    emax, _, mism = O.column_error(got[:, 1:], want[:, 1:])   # f = 0: NaN in both
    print(f"mode {mode}: emax={emax:.2e}")
    assert mism == 0
    assert emax <= TOL


@pytest.fixture(scope="module")
def single():
    from ccfj import synth
    return synth.noise_single()


def test_noise_single_full_size_vs_oracle(ccfj, single):
    """300 pairs x 512 f x 500 c, itype 1, func 0 through fj_noise, against
    the reference's Python layer restated over the FP64 oracle."""
    d = single
    got = ccfj.fj_noise(d["uf"], d["r"], d["c"], d["f"], itype=1, func=0)
    assert got.dtype == np.float32 and got.shape == (500, 512)
    want = O.fj_noise(d["uf"], d["r"], d["c"], d["f"], itype=1, func=0, engine="fp64")
    emax, erms, mism = O.column_error(got, want)
    print(f"noise_single itype1 func0: emax={emax:.2e} erms={erms:.2e}")
    assert mism == 0 and emax <= TOL
    # dispersion ridge: the transform peaks at the model phase velocity
    c = d["c"]
    pick = c[np.argmax(got, axis=0)]
    band = (d["f"] > 0.05) & (d["f"] < 0.45)
    assert np.median(np.abs(pick[band] - d["c0"][band])) < 2 * (c[1] - c[0]) * 5


def test_noise_single_vs_fjcpu(ccfj, single):
    d = single
    got = ccfj.fj_noise(d["uf"], d["r"], d["c"], d["f"], itype=1, func=0)
    fjcpu = O.fj_noise(d["uf"], d["r"], d["c"], d["f"], itype=1, func=0, engine="port")
    ref64 = O.fj_noise(d["uf"], d["r"], d["c"], d["f"], itype=1, func=0, engine="fp64")
    e_ref = O.column_error(fjcpu, ref64)[0]
    e_cmp = O.column_error(got, fjcpu)[0]
    print(f"FJcpu vs oracle {e_ref:.2e}; GPU vs FJcpu {e_cmp:.2e}")
    assert e_cmp <= 2 * e_ref + TOL


@pytest.mark.parametrize("itype", [0, 1])
@pytest.mark.parametrize("func", [0, 1])
def test_fj_noise_wrapper_semantics(ccfj, single, itype, func):
    d = single
    sub = slice(0, 120)
    uf, r = d["uf"][sub, ::8], d["r"][sub]
    perm = np.random.default_rng(7).permutation(len(r))   # unsorted input
    f = d["f"][::8]
    got = ccfj.fj_noise(uf[perm], r[perm], d["c"][::5], f, itype=itype, func=func)
    want = O.fj_noise(uf, r, d["c"][::5], f, itype=itype, func=func, engine="fp64")
    emax, _, mism = O.column_error(got, want)
    assert mism == 0 and emax <= TOL, emax
    assert np.allclose(np.nanmax(np.abs(got), axis=0), 1.0)


def test_fj_noise_fstride(ccfj, single):
    d = single
    uf, f = d["uf"][:60], d["f"][::3]
    got = ccfj.fj_noise(uf, d["r"][:60], d["c"][::10], f, fstride=3, itype=0)
    want = O.fj_noise(uf, d["r"][:60], d["c"][::10], f, fstride=3, itype=0, engine="fp64")
    assert O.column_error(got, want)[0] <= TOL


# --------------------------------------------------------- earthquake / MWFJ
def _quake(seed=3, nr=48, npts=1024, Fs=20.0):
    rng = np.random.default_rng(seed)
    r = np.sort(rng.uniform(2e3, 60e3, nr))
    t = np.arange(npts) / Fs
    u = np.zeros((nr, npts))
    for cm, a in ((3200.0, 1.0), (3800.0, 0.4)):
        arr = r / cm
        u += a * np.exp(-((t[None, :] - 5 - arr[:, None]) * 2.0) ** 2) * np.cos(2 * np.pi * 1.0 * (t[None, :] - arr[:, None]))
    u += 0.05 * rng.standard_normal(u.shape)
    f = np.arange(1, 120) * Fs / npts
    c = np.linspace(2000.0, 5000.0, 61)
    return u.astype(np.float32), r, c, f, Fs


@pytest.mark.parametrize("itype", [0, 1])
@pytest.mark.parametrize("func", [0, 1])
def test_fj_earthquake_wrapper(ccfj, itype, func):
    u, r, c, f, _ = _quake()
    perm = np.random.default_rng(1).permutation(len(r))
    got = ccfj.fj_earthquake(u[perm], r[perm], c, f, itype=itype, func=func)
    want = O.fj_earthquake(u, r, c, f, itype=itype, func=func, engine="fp64")
    emax, _, mism = O.column_error(got, want)
    assert mism == 0 and emax <= TOL, emax


@pytest.mark.parametrize("itype", [0, 1])
@pytest.mark.parametrize("func", [0, 1])
def test_mwfj_wrapper(ccfj, itype, func):
    """Device MWFJ (tapering, double rFFT, windows-as-right-hand-sides MFMA
    for itype 0) against the reference's per-window loop restated over the
    FP64 oracle; includes a reversed and an out-of-range window."""
    u, r, c, f, Fs = _quake(nr=24, npts=512)
    nwin = 5
    v = np.array([2500.0, 3000.0, 3500.0, 4500.0, 5500.0, 6500.0])
    winl = np.stack([r / v[w + 1] for w in range(nwin)])
    winr = np.stack([r / v[w] + 8.0 for w in range(nwin)])
    winl[1, 3], winr[1, 3] = 20.0, 10.0       # reversed: all zero
    winr[2, 5] = 1e3                          # beyond the trace end
    got = ccfj.MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=itype, func=func)
    want = O.MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=itype, func=func,
                  engine="fp64")
    assert got.dtype == np.float64 and got.shape == (nwin, len(c), len(f))
    for w in range(nwin):
        emax, _, mism = O.column_error(got[w], want[w])
        assert mism == 0 and emax <= TOL, (w, emax)


@pytest.mark.parametrize("func", [0, 1])
def test_mwfj_many_receivers(ccfj, func):
    """More receivers (1600) than the contraction kernel stages in LDS
    (1536 rows): its station arrays are then read from global memory."""
    u, r, c, f, Fs = _quake(nr=1600, npts=256)
    nwin = 2
    winl = np.stack([r / 4000.0, r / 3000.0])
    winr = np.stack([r / 2500.0 + 8.0, r / 2000.0 + 8.0])
    got = ccfj.MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=0, func=func)
    want = O.MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=0, func=func,
                  engine="fp64")
    for w in range(nwin):
        emax, _, mism = O.column_error(got[w], want[w])
        assert mism == 0 and emax <= TOL, (w, emax)


@pytest.mark.parametrize("npts", [511, 512])
def test_mwfj_trace_length_parity(ccfj, npts):
    """Odd trace lengths take the full rFFT (D2Z), even ones the half-length
    complex FFT with the used bins formed on read; both against the oracle."""
    u, r, c, f, Fs = _quake(nr=20, npts=npts, seed=5)
    f = f[:100]
    nwin = 3
    winl = np.stack([r / 4500.0 + w for w in range(nwin)])
    winr = winl + 9.0
    got = ccfj.MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=0, func=1)
    want = O.MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=0, func=1, engine="fp64")
    for w in range(nwin):
        emax, _, mism = O.column_error(got[w], want[w])
        assert mism == 0 and emax <= TOL, (w, emax)


def test_mwfj_full_size_sampled(ccfj):
    """MWFJ at the active-source size (480 receivers x 16,384 samples, 64
    windows, 820 f, 2000 c) on the GPU; three windows and five frequency
    columns against the FP64 oracle (the reference's window, numpy's rFFT,
    the FP64 trapezoid sum, |Re + i Im|, per-column normalisation)."""
    from ccfj import synth
    d = synth.active_source(seed=2)
    u, r, f, c, Fs = d["u"], d["r"], d["f"], d["c"], d["Fs"]
    nr, npts, nwin = len(r), u.shape[1], d["nwin"]
    got = ccfj.MWFJ(u, r, c, f, Fs, nwin, d["winl"], d["winr"], 0.9, 1, 0, 0)
    assert got.shape == (nwin, len(c), len(f)) and got.dtype == np.float64
    cols = np.array([1, 100, 300, 550, len(f) - 1])
    idx = np.argsort(r)
    for w in (0, 31, nwin - 1):
        tap = np.stack([O.win(npts, Fs, d["winl"][w, j], d["winr"][w, j], 0.9) for j in range(nr)])
        spec = np.fft.rfft(u.astype(np.float64) * tap, axis=1)[idx][:, :len(f)]
        re = O.oracle64(np.ascontiguousarray(spec.real, np.float32), r[idx], f, c, 0, 0, fidx=cols)
        im = O.oracle64(np.ascontiguousarray(spec.imag, np.float32), r[idx], f, c, 0, 0, fidx=cols)
        want = np.sqrt(re.astype(np.float64) ** 2 + im.astype(np.float64) ** 2)
        want = want / np.max(want, axis=0, keepdims=True)
        emax, erms, mism = O.column_error(got[w][:, cols], want)
        print(f"MWFJ window {w}: emax={emax:.2e} erms={erms:.2e}")
        assert mism == 0 and emax <= TOL, (w, emax)


def test_mwfj_many_windows_mfma_tiles(ccfj):
    """More than 64 windows: the right-hand sides span two MFMA groups."""
    u, r, c, f, Fs = _quake(nr=16, npts=256, seed=9)
    nwin = 70
    rng = np.random.default_rng(4)
    t0 = rng.uniform(0, 6, (nwin, 1)) + r[None, :] / 4000.0
    winl, winr = t0, t0 + rng.uniform(2, 6, (nwin, 1))
    got = ccfj.MWFJ(u, r, c[::4], f[::3], Fs, nwin, winl, winr, taper=0.8, itype=0)
    want = O.MWFJ(u, r, c[::4], f[::3], Fs, nwin, winl, winr, taper=0.8, itype=0, engine="fp64")
    for w in (0, 31, 32, 63, 64, 69):
        emax, _, mism = O.column_error(got[w], want[w])
        assert mism == 0 and emax <= TOL, (w, emax)


@pytest.mark.parametrize("itype", [0, 1])
@pytest.mark.parametrize("func", [0, 1])
def test_mwfj_goertzel_and_fft_spectra(ccfj, monkeypatch, itype, func):
    """The used bins come from the Goertzel recurrence over each window's
    support, or from the FFT of the whole tapered rows, whichever costs less
    (CCFJ_MWFJ_SPECTRA forces one).  Both paths against the oracle, and
    against each other, on windows of every kind: short, tapered, reaching
    past the end of the trace, empty, and degenerate ones that travel as
    explicit rows."""
    u, r, c, f, Fs = _quake(nr=32, npts=512, seed=6)
    nwin = 6
    t0 = np.linspace(2.0, 9.0, nwin)[:, None] + r[None, :] / 4000.0
    width = np.array([0.5, 2.0, 6.0, 40.0, 0.0, 0.01])[:, None] + 0.0 * r[None, :]
    winl, winr = t0, t0 + width
    winl[3] = 0.0   # the whole trace and beyond
    out = {}
    for path in ("goertzel", "fft"):
        monkeypatch.setenv("CCFJ_MWFJ_SPECTRA", path)
        out[path] = ccfj.MWFJ(u, r, c, f[::2], Fs, nwin, winl, winr, taper=0.6, itype=itype,
                              func=func)
    monkeypatch.delenv("CCFJ_MWFJ_SPECTRA")
    want = O.MWFJ(u, r, c, f[::2], Fs, nwin, winl, winr, taper=0.6, itype=itype, func=func,
                  engine="fp64")
    for w in range(nwin):
        for path in ("goertzel", "fft"):
            emax, _, mism = O.column_error(out[path][w], want[w])
            assert mism == 0 and emax <= TOL, (path, w, emax)
        emax, _, mism = O.column_error(out["goertzel"][w], out["fft"][w])
        assert mism == 0 and emax <= 1e-6, (w, emax)


@pytest.mark.parametrize("func", [0, 1])
def test_mwfj_zero_frequency_and_nonfinite_coefficients(ccfj, func):
    """The contraction takes fp16 split products where a frequency's scale
    allows and FP32 products where it does not: at f = 0 (entries sqrt(r) Z0(0),
    -inf for the Y basis) and where a coefficient is not finite.  Every
    column, f = 0 included, against the oracle; a NaN sample in one trace
    makes every output NaN, as in the reference."""
    u, r, c, f, Fs = _quake(nr=40, npts=256, seed=3)
    f = np.concatenate([[0.0], f[:40]])
    nwin = 3
    winl = np.stack([r / 4500.0 + w for w in range(nwin)])
    winr = winl + 12.0
    got = ccfj.MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=0, func=func)
    want = O.MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=0, func=func, engine="fp64")
    for w in range(nwin):
        emax, _, mism = O.column_error(got[w], want[w])
        assert mism == 0 and emax <= TOL, (w, emax)
    bad = u.copy()
    bad[7, 100] = np.nan
    got = ccfj.MWFJ(bad, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=0, func=func)
    want = O.MWFJ(bad, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=0, func=func,
                  engine="fp64")
    assert np.array_equal(np.isnan(got), np.isnan(want))
    assert np.isnan(got).all()


@pytest.mark.parametrize("itype", [0, 1])
def test_mwfj_device_pointers_match_host_entry(ccfj, ctx, itype):
    """ccfj_mwfj_dev on device arrays gives the host entry point's result
    bit for bit, on two calls in a row (the second smaller, reusing the
    context's work arrays); a bad xidx is an error on the host entry."""
    import torch
    dev = torch.device("cuda:0")
    for nr, npts, nwin in ((24, 512, 5), (16, 256, 3)):
        u, r, c, f, Fs = _quake(nr=nr, npts=npts)
        v = np.array([2500.0, 3000.0, 3500.0, 4500.0, 5500.0, 6500.0])
        winl = np.stack([r / v[w + 1] for w in range(nwin)])
        winr = np.stack([r / v[w] + 8.0 for w in range(nwin)])
        winl[1, 3], winr[1, 3] = 20.0, 10.0       # reversed: an explicit row
        indx = np.argsort(r)
        us = np.ascontiguousarray(u[indx], np.float64)
        rs = r[indx].astype(np.float32)
        f32, c32 = f.astype(np.float32), c.astype(np.float32)
        n1, n2, dl, xidx, xrows = ccfj._window_descriptions(winl, winr, indx, npts, Fs, 0.9)
        assert len(xrows) >= 1
        want = ctx.mwfj(itype, 0, 1, us, rs, f32, c32, n1, n2, dl, xidx, xrows)
        t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev)
             for a in (us, rs, f32, c32, n1, n2, dl, xidx, xrows)]
        out = torch.full((nwin, len(c), len(f)), float("nan"), dtype=torch.float64, device=dev)
        st = torch.cuda.Stream(device=dev)
        ctx.mwfj_dev(itype, 0, 1, nr, npts, nwin, len(f), len(c), *[a.data_ptr() for a in t],
                     len(xrows), out.data_ptr(), st.cuda_stream)
        st.synchronize()
        assert np.array_equal(out.cpu().numpy(), want)
    bad = xidx.copy()
    bad[0, 0] = len(xrows)
    with pytest.raises(RuntimeError, match="xidx"):
        ctx.mwfj(itype, 0, 1, us, rs, f32, c32, n1, n2, dl, bad, xrows)


def test_threads_share_the_default_context(ccfj):
    """Python threads calling the drop-in API at once (the GIL is released
    in the native call) get the results of serial calls: host-pointer entry
    points hold the context's lock."""
    import threading
    u, r, c, f, _ = _quake(nr=40, npts=512, seed=11)
    spec = np.fft.rfft(u, axis=1)[:, :100].astype(np.complex64)
    jobs = [(itype, func) for itype in (0, 1) for func in (0, 1)] * 2
    serial = {j: ccfj.fj_noise(spec, r, c, f[:100], itype=j[0], func=j[1]) for j in jobs[:4]}
    got = {}

    def work(k, j):
        got[k] = ccfj.fj_noise(spec, r, c, f[:100], itype=j[0], func=j[1])

    ts = [threading.Thread(target=work, args=(k, j)) for k, j in enumerate(jobs)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    for k, j in enumerate(jobs):
        assert np.array_equal(got[k], serial[j]), (k, j)


# ------------------------------------------------------------- edge cases
def test_edges_and_errors(ccfj, ctx):
    r = np.array([1000.0, 1000.0, 1000.05, 3000.0, 9000.0], np.float32)
    U = np.random.default_rng(0).standard_normal((5, 11)).astype(np.float32)
    f = np.concatenate([[0.0], np.linspace(0.05, 1.0, 10)]).astype(np.float32)
    c = np.linspace(800, 4000, 7).astype(np.float32)
    for itype in (0, 1):
        for mode, yb in ((ccfj.MODE_J, 0), (ccfj.MODE_Y, 1)):
            got = native(ctx, mode, itype, U, r, f, c)
            want = O.oracle64(U, r, f, c, itype, yb)
            emax, _, mism = O.column_error(got, want)
            assert mism == 0 and emax <= TOL, (itype, mode, emax)
    # no stations / one station: zeros, as the reference's empty loops
    # (also at f = 0, where a segment would divide by k = 0)
    for itype in (0, 1):
        for mode in (ccfj.MODE_J, ccfj.MODE_Y):
            assert np.all(native(ctx, mode, itype, U[:1], r[:1], f, c) == 0)
    z = ctx.transform(ccfj.MODE_J, 0, False, np.zeros((0, 11), np.float32), None,
                      np.zeros(0, np.float32), f, c)
    assert z.shape == (7, 11) and np.all(z == 0)
    # itype outside {0,1}: the reference leaves zeros -> NaN after normalising
    with np.errstate(invalid="ignore"):
        assert np.all(np.isnan(ccfj.fj_noise(U.T.copy().T, r.astype(np.float64), c, f, itype=2)))
    # func outside {0,1}: prints and returns 0
    assert ccfj.fj_noise(U, r, c, f, func=2) == 0
    # fj_earthquake with fstride > 1 strides twice and fails in reshape
    with pytest.raises(ValueError):
        ccfj.fj_earthquake(np.ones((5, 64)), r, c, f[:10], fstride=2)
    # C-ABI argument errors surface as RuntimeError with the library's text
    with pytest.raises(RuntimeError, match="mode"):
        ctx.transform(17, 0, False, U, None, r, f, c)


def test_reference_compatible_entry_points(ccfj, ctx, golden):
    from ccfj import _ccfj
    U, r, f, c = golden_case(golden, "branches")
    nr, nf, nc = len(r), len(f), len(c)
    for itype in (0, 1):
        for fn, mode in ((_ccfj.fj_full, ccfj.MODE_J), (_ccfj.fHr_full, ccfj.MODE_J),
                         (_ccfj.fHi_full, ccfj.MODE_Y)):
            out = np.zeros(nf * nc, np.float32)
            fn(U.reshape(-1).copy(), r.copy(), f.copy(), out, c.copy(), nc, nr, nf, itype)
            assert np.array_equal(out.reshape(nc, nf), native(ctx, mode, itype, U, r, f, c))
    out = np.zeros(nf * nc, np.float32)
    with pytest.raises(RuntimeError):
        _ccfj.fj_full(U.reshape(-1).copy(), r.copy(), f.copy(), out, c.copy(), nc, nr, nf, 3)


def test_plain_fp64_baseline(ccfj, golden):
    from ccfj import _ccfj
    U, r, f, c = golden_case(golden, "cfg1s")
    for yb in (0, 1):
        for itype in (0, 1):
            out, ms = _ccfj.plain_fp64(0, yb, itype, U, r, f, c)
            emax, _, mism = O.column_error(out, O.oracle64(U, r, f, c, itype, yb))
            # FP64 throughout, but the Zhang-Jin Struve polynomials carry
            # ~1e-8 relative error, amplified by the closed form's cancellation
            assert mism == 0 and emax < TOL, (yb, itype, emax)
            assert ms > 0


def test_alloc_failure_is_reported_not_fatal():
    code = ("import sys; sys.path.insert(0, %r); import numpy as np, ccfj\n"
            "try:\n"
            "    ccfj.fj(np.ones((3, 4), np.float32), np.arange(3.) * 100, np.arange(1., 6.) * 1e3, np.arange(4.) * .1)\n"
            "    print('NO-ERROR')\n"
            "except RuntimeError as e:\n"
            "    print('ERR', e)\n" % f"{ROOT}/cc-fjpy_amd")
    p = subprocess.run([sys.executable, "-c", code], env={**__import__("os").environ, "CCFJ_FAIL_ALLOC_AT": "1"},
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "ERR" in p.stdout and "injected" in p.stdout


# ------------------------------------------------------- dense-array workload
@pytest.fixture(scope="module")
def dense():
    from ccfj import synth
    d = synth.noise_dense(real_only=True)
    idx = np.argsort(d["r"])
    return dict(U=np.ascontiguousarray(d["uf"][idx]), r=d["r"][idx].astype(np.float32),
                f=d["f"].astype(np.float32), c=d["c"].astype(np.float32))


# eight columns stratified over 0.005-2 Hz, the last (2 Hz, the largest
# arguments k r) included
DENSE_COLS = [7, 600, 1200, 1800, 2400, 3000, 3600, 4095]


@pytest.fixture(scope="module")
def dense_oracle(ccfj, dense):
    """FP64 oracle at DENSE_COLS: J0 basis of U and Y0 basis of the
    Hilbert-transformed U, per itype (computed once, shared by func 0/1)."""
    d = dense
    H = ccfj.hilbert_rows(d["U"])
    cache = {"H": H}
    for itype in (0, 1):
        print(f"dense oracle itype={itype}", flush=True)
        cache[itype, 0] = O.oracle64(d["U"], d["r"], d["f"], d["c"], itype, 0, fidx=DENSE_COLS)
        cache[itype, 1] = O.oracle64(H, d["r"], d["f"], d["c"], itype, 1, fidx=DENSE_COLS)
    return cache


@pytest.mark.parametrize("itype", [0, 1])
@pytest.mark.parametrize("func", [0, 1])
def test_dense_full_size_sampled_columns(ccfj, ctx, dense, dense_oracle, itype, func):
    """19,900 pairs x 4096 f x 3000 c at full size on the GPU; eight
    frequency columns from 0.005 to 2 Hz checked against the FP64 oracle."""
    d = dense
    if func == 0:
        got = native(ctx, ccfj.MODE_J, itype, d["U"], d["r"], d["f"], d["c"])
        want = dense_oracle[itype, 0]
    else:
        got = native(ctx, ccfj.MODE_J_MINUS_Y, itype, d["U"], d["r"], d["f"], d["c"],
                     U1=dense_oracle["H"])
        want = dense_oracle[itype, 0] - dense_oracle[itype, 1]
    emax, erms, mism = O.column_error(got[:, DENSE_COLS], want)
    print(f"dense itype={itype} func={func}: emax={emax:.2e} erms={erms:.2e}")
    assert mism == 0 and emax <= TOL
    assert np.all(np.isfinite(got))


def test_dense_linearity(ccfj, ctx, dense):
    """Size-independent property at full size: F[U + 2V] = F[U] + 2 F[V]."""
    d = dense
    V = np.roll(d["U"], 1, axis=1)
    a = native(ctx, ccfj.MODE_J, 1, d["U"] + 2 * V, d["r"], d["f"], d["c"]).astype(np.float64)
    b = native(ctx, ccfj.MODE_J, 1, d["U"], d["r"], d["f"], d["c"]).astype(np.float64)
    b += 2 * native(ctx, ccfj.MODE_J, 1, V, d["r"], d["f"], d["c"])
    assert O.column_error(a, b)[0] < 2 * TOL
