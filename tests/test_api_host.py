"""Host-side logic of the ccfj package (no GPU): the window, the Hilbert
transform along frequency, station pairs, staging errors and the
reference's behaviour for bad func / itype."""
import io
import contextlib
import warnings

import numpy as np
import pytest
from scipy import fftpack

import ccfj
from oracle import oracle as O


def test_win_matches_reference_semantics():
    """Vectorised window == the element-by-element restatement of
    src/ccfj.pyx:148-158, including degenerate windows (reversed, out of
    range, negative ends) that use Python's negative indexing."""
    rng = np.random.default_rng(0)
    for _ in range(400):
        npts = int(rng.integers(1, 400))
        Fs = float(rng.choice([1.0, 4.0, 100.0]))
        T1 = float(rng.uniform(-2, 1.2) * npts / Fs)
        T2 = float(rng.uniform(-0.5, 1.5) * npts / Fs)
        taper = float(rng.choice([0.9, 0.8, 0.5, 0.0, 1.0]))
        got = ccfj.win(npts, Fs, T1, T2, taper)
        want = O.win(npts, Fs, T1, T2, taper)
        assert np.array_equal(got, want), (npts, Fs, T1, T2, taper)


@pytest.mark.parametrize("n", [64, 63, 4096, 1])
def test_hilbert_rows_matches_fftpack(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((7, n)).astype(np.float32)
    want = np.array([fftpack.hilbert(row) for row in x]).astype(np.float32)
    got = ccfj.hilbert_rows(x)
    assert got.dtype == np.float32
    assert np.max(np.abs(got - want)) <= 2e-7 * max(1.0, np.abs(want).max())
    # complex rows: the reference stores hilbert(Re) after the float32 cast
    z = (x + 1j * rng.standard_normal(x.shape)).astype(np.complex64)
    want_c = np.array([np.real(fftpack.hilbert(row)) for row in z]).astype(np.float32)
    assert np.max(np.abs(ccfj.hilbert_rows(z) - want_c)) <= 2e-7 * max(1.0, np.abs(want_c).max())


def test_hilbert_of_cosine_is_minus_sine():
    t = np.arange(256) / 256
    y = ccfj.hilbert_rows(np.cos(2 * np.pi * 5 * t)[None, :])[0]
    assert np.max(np.abs(y + np.sin(2 * np.pi * 5 * t))) < 1e-6


def test_station_pairs():
    for n in (0, 1, 2, 5, 30):
        got = ccfj.GetStationPairs(n)
        want = [v for ii in range(n) for jj in range(ii + 1, n) for v in (ii, jj)]
        assert got.dtype == np.int32 and got.tolist() == want


def test_bad_func_prints_and_returns_zero():
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        assert ccfj.fj_noise(np.ones((3, 4)), np.arange(3.), np.arange(1., 3.), np.arange(4.), func=5) == 0
    assert "set func as 0" in buf.getvalue()


def test_bad_itype_gives_nan_without_touching_the_gpu():
    with np.errstate(invalid="ignore"):
        out = ccfj.fj_noise(np.ones((3, 4)), np.arange(3.), np.arange(1., 3.), np.arange(4.), itype=7)
    assert out.shape == (2, 4) and np.all(np.isnan(out))


def test_staging_shape_error_like_reference():
    # ceil(ncol / fstride) != len(f): the reference's reshape raises
    with pytest.raises(ValueError):
        ccfj.fj(np.ones((3, 10), np.float32), np.arange(3.), np.arange(1., 3.), np.arange(4.), fstride=2)


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    from conftest import ROOT
    p = subprocess.run([sys.executable, f"{ROOT}/bench.py", "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr


def test_gather_real_rows_is_sort_then_cast():
    """The threaded row gather fj_noise uses equals the reference's
    uf[indx] followed by the float32 cast, for complex64 and float32 input."""
    from ccfj import _ccfj
    rng = np.random.default_rng(3)
    for dt in (np.complex64, np.float32):
        uf = rng.standard_normal((37, 129)).astype(np.float32)
        if dt == np.complex64:
            uf = (uf + 1j * rng.standard_normal((37, 129))).astype(np.complex64)
        indx = np.argsort(rng.uniform(size=37))
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            want = np.array(uf[indx], dtype=np.float32)
        assert np.array_equal(_ccfj.gather_real_rows(uf, indx), want)
    with pytest.raises(ValueError):
        _ccfj.gather_real_rows(uf, np.arange(36))
    with pytest.raises(TypeError):
        _ccfj.gather_real_rows(uf.astype(np.float64), indx)
