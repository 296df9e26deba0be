"""The drop-in boundary on the GPU: the reference's un-normalised helpers and
module-level native shims, the real-data flow of its noise example, the
func = 1 stride ordering and odd row lengths of the device Hilbert
transform, float32 window edges in MWFJ, and two contexts running at once.
Bars as in test_gpu_parity: |X - X*| / max_c |X*[:, f]| <= 1e-5 against the
FP64 oracle, identical NaN patterns."""
import threading
import warnings

import numpy as np
import pytest

from oracle import cc_ref as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def single():
    from ccfj import synth
    return synth.noise_single()


def _quiet(fn, *a, **k):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return fn(*a, **k)


@pytest.mark.parametrize("itype", [0, 1])
def test_unnormalised_helpers_vs_oracle(ccfj, single, itype):
    """fj / fhr (J0 basis) and fh / fhi (Y0 basis) stage uf[:, ::fstride] as
    float32 and return the raw float32 [nc, nf] transform, no normalisation
    (src/ccfj.pyx:161-198, src/ccfj-gpu.pyx:181-239)."""
    d = single
    uf, r = d["uf"][:150, :200], d["r"][:150]
    f, c = d["f"][:200:2], d["c"][::4]
    U = np.ascontiguousarray(np.real(uf[:, ::2]), np.float32)
    J = O.oracle64(U, r, f, c, itype, 0)
    Y = O.oracle64(U, r, f, c, itype, 1)
    for fn, want in ((ccfj.fj, J), (ccfj.fhr, J), (ccfj.fh, Y), (ccfj.fhi, Y)):
        got = _quiet(fn, uf, r, c, f, fstride=2, itype=itype)
        assert got.dtype == np.float32 and got.shape == (len(c), len(f))
        emax, erms, mism = O.column_error(got, want)
        print(f"{fn.__name__} itype={itype}: emax={emax:.2e}")
        assert mism == 0 and emax <= TOL, (fn.__name__, emax)
        assert not np.allclose(np.max(np.abs(got), axis=0), 1.0)   # not normalised
    # the CPU build's positional nThread
    got = _quiet(ccfj.fj, uf, r, c, f, 2, itype, 20)
    assert O.column_error(got, J)[0] <= TOL


def test_module_level_native_shims(ccfj, single):
    """ccfj.fj_full (with and without the CPU build's nThread), fH_full,
    fHr_full, fHi_full on flat float32 buffers, as src/ccfj.pyx:72-113 and
    src/ccfj-gpu.pyx:75-135 expose them."""
    d = single
    U = np.ascontiguousarray(np.real(d["uf"][:100, :64]), np.float32)
    r = np.asarray(d["r"][:100], np.float32)
    f = np.asarray(d["f"][:64], np.float32)
    c = np.asarray(d["c"][::5], np.float32)
    nr, nf, nc = len(r), len(f), len(c)
    for itype in (0, 1):
        J = O.oracle64(U, r, f, c, itype, 0)
        Y = O.oracle64(U, r, f, c, itype, 1)
        calls = [(lambda o: ccfj.fj_full(U.reshape(-1), r, f, o, c, nc, nr, nf, itype), J),
                 (lambda o: ccfj.fj_full(U.reshape(-1), r, f, o, c, nc, nr, nf, itype, 20), J),
                 (lambda o: ccfj.fHr_full(U.reshape(-1), r, f, o, c, nc, nr, nf, itype), J),
                 (lambda o: ccfj.fH_full(U.reshape(-1), r, f, o, c, nc, nr, nf, itype), Y),
                 (lambda o: ccfj.fH_full(U.reshape(-1), r, f, o, c, nc, nr, nf, itype, 8), Y),
                 (lambda o: ccfj.fHi_full(U.reshape(-1), r, f, o, c, nc, nr, nf, itype), Y)]
        for k, (call, want) in enumerate(calls):
            out = np.zeros(nf * nc, np.float32)
            assert call(out) is None
            emax, _, mism = O.column_error(out.reshape(nc, nf), want)
            assert mism == 0 and emax <= TOL, (itype, k, emax)
    with pytest.raises(ValueError):     # a float64 buffer does not match float[:]
        ccfj.fj_full(U.reshape(-1).astype(np.float64), r, f, np.zeros(nf * nc, np.float32), c,
                     nc, nr, nf, 0)


def test_cc_full_matches_cc(ccfj):
    """ccfj.CC_full fills ncfsr / ncfsi as the reference's CC does before it
    combines them (src/ccfj.pyx:44-68, :125-146)."""
    rng = np.random.default_rng(12)
    nsta, npts, fftlen, nf = 5, 4000, 512, 120
    x = rng.normal(size=(nsta, npts)).astype(np.float32)
    se = np.tile(np.array([0, npts], np.int32), nsta)
    se[3] = npts // 2
    pairs = ccfj.GetStationPairs(nsta)
    npairs = len(pairs) // 2
    re = np.zeros(nf * npairs, np.float32)
    im = np.zeros(nf * npairs, np.float32)
    data = x.reshape(-1).copy()
    ccfj.CC_full(npts, nsta, nf, 1, fftlen, fftlen // 2, 0, 1, 8, npairs, pairs, se, data, re, im)
    want, _ = C.cross_correlate(x.reshape(-1), nsta, npts, nf, fftlen, fftlen // 2, pairs, se)
    got = (re + 1j * im).reshape(npairs, nf)
    assert np.max(np.abs(got - want)) <= TOL * np.max(np.abs(want))
    cc = ccfj.CC(npts, nsta, nf, fftlen, pairs, se, x.reshape(-1).copy(), overlaprate=0.5)
    assert np.max(np.abs(cc - got)) <= TOL * np.max(np.abs(want))


def test_example_noise_flow(ccfj, tmp_path):
    """examples/example_noise.py:6-16: summed.npz -> r * 1e3 -> fj_noise on
    np.real(ncfs) (a strided float32 view) for all four variants, with
    c = linspace(2000, 5000, 1000)."""
    from ccfj import io, synth
    d = synth.noise_single(seed=9, nr=160, nf=96, nc=10)
    perm = np.random.default_rng(1).permutation(160)
    io.save_summed(tmp_path / "summed.npz", d["uf"][perm], d["r"][perm] / 1e3, d["f"])
    ncfs, r, f = io.load_summed(tmp_path / "summed.npz")
    assert ncfs.dtype == np.complex64 and np.all(np.diff(r) >= 0)
    c = np.linspace(2000, 5000, 1000)
    re = np.real(ncfs)
    assert not re.flags.c_contiguous
    for itype in (0, 1):
        for func in (0, 1):
            got = ccfj.fj_noise(re, r, c, f, fstride=1, itype=itype, func=func)
            want = O.fj_noise(re, r, c, f, fstride=1, itype=itype, func=func, engine="fp64")
            emax, erms, mism = O.column_error(got, want)
            print(f"example flow itype={itype} func={func}: emax={emax:.2e}")
            assert got.shape == (1000, len(f)) and got.dtype == np.float32
            assert mism == 0 and emax <= TOL, (itype, func, emax)


@pytest.mark.parametrize("itype", [0, 1])
@pytest.mark.parametrize("ncol,fstride", [(120, 3), (121, 3), (121, 1), (64, 1)])
def test_fj_noise_hankel_stride_and_row_length(ccfj, single, itype, ncol, fstride):
    """func 1 takes the Hilbert transform of the full rows (even or odd
    length: the Nyquist bin exists only for even lengths) and then the
    column stride (src/ccfj.pyx:215-220), on unsorted input."""
    d = single
    uf, r = d["uf"][:70, :ncol], d["r"][:70]
    f = d["f"][:ncol][::fstride]
    c = d["c"][::10]
    perm = np.random.default_rng(ncol).permutation(70)
    got = _quiet(ccfj.fj_noise, uf[perm], r[perm], c, f, fstride=fstride, itype=itype, func=1)
    want = _quiet(O.fj_noise, uf, r, c, f, fstride=fstride, itype=itype, func=1, engine="fp64")
    emax, _, mism = O.column_error(got, want)
    print(f"hankel ncol={ncol} fstride={fstride} itype={itype}: emax={emax:.2e}")
    assert mism == 0 and emax <= TOL


def test_mwfj_float32_window_edges(ccfj):
    """float32 winl / winr: the window sample counts follow the reference's
    float32 arithmetic (src/ccfj.pyx:149-152), here checked end to end
    against the restated per-window loop."""
    rng = np.random.default_rng(3)
    nr, npts, Fs = 16, 600, 100.0
    r = np.sort(rng.uniform(2e3, 30e3, nr))
    u = rng.standard_normal((nr, npts)).astype(np.float32)
    f = np.arange(1, 60) * Fs / npts
    c = np.linspace(1500.0, 4500.0, 50)
    nwin = 3
    winl = np.stack([r / 4000.0 + 0.29 * w for w in range(nwin)]).astype(np.float32)
    winr = (winl + np.float32(3.0)).astype(np.float32)
    got = ccfj.MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=0)
    want = O.MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, itype=0, engine="fp64")
    for w in range(nwin):
        emax, _, mism = O.column_error(got[w], want[w])
        assert mism == 0 and emax <= TOL, (w, emax)


def test_two_contexts_in_two_threads(ccfj, single):
    """Two contexts on one device, each driven by its own thread, run
    fj_noise func = 1 (device Hilbert transform through hipFFT) on the same
    sizes at once; every result equals the serial one bit for bit (each
    context has its own FFT plans and work areas)."""
    from ccfj import _ccfj
    d = single
    idx = np.argsort(d["r"])
    U = np.ascontiguousarray(np.real(d["uf"][idx]), np.float32)
    r = np.asarray(d["r"][idx], np.float32)
    f = np.asarray(d["f"], np.float32)
    c = np.asarray(d["c"][::2], np.float32)
    ctxs = [_ccfj.Context(0), _ccfj.Context(0)]
    serial = {it: ctxs[0].fj_noise(it, 1, 1, U, r, f, c) for it in (0, 1)}
    got, errs = [], []

    def work(ctx, it):
        try:
            for _ in range(4):
                got.append((it, ctx.fj_noise(it, 1, 1, U, r, f, c)))
        except Exception as exc:   # surfaced below
            errs.append(exc)

    ts = [threading.Thread(target=work, args=(ctxs[k], k)) for k in (0, 1)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    assert len(got) == 8
    for it, out in got:
        assert np.array_equal(out.view(np.uint32), serial[it].view(np.uint32)), it
