"""The CPU restatement of the cross-correlation (oracle/cc_ref.py), pinned
by the cross-correlation theorem against a brute-force time-domain
correlation, by the reference's window rules, and by whitening's unit
magnitude."""
import numpy as np

from oracle import cc_ref as C


def test_window_count_rule():
    assert C.windows(1000, 1000, 100) == (1000, 1)
    assert C.windows(1000, 5000, 100) == (1000, 1)
    assert C.windows(1000, 256, 128) == (256, 1 + (1000 - 256) // 128)
    assert C.windows(1000, 256, 77) == (256, 1 + 744 // 77)


def test_window_use_rule():
    se = np.array([0, 1000, 100, 1000, 0, 700], np.int32)
    # window k covers [k*step, k*step + fftlen): needs start <= k*step and
    # k*step + fftlen < end
    assert list(C.used(se, 0, 128, 256)) == [True, False, True]
    assert list(C.used(se, 1, 128, 256)) == [True, True, True]
    assert list(C.used(se, 3, 128, 256)) == [True, True, True]    # 384 + 256 = 640 < 700
    assert list(C.used(se, 4, 128, 256)) == [True, True, False]   # 512 + 256 = 768 >= 700


def test_single_window_is_circular_cross_correlation():
    rng = np.random.default_rng(0)
    n = 64
    x = rng.normal(size=(2, n)).astype(np.float32)
    nf = n // 2 + 1
    out, count = C.cross_correlate(x.reshape(-1), 2, n, nf, n, n, [0, 1], [0, n + 1, 0, n + 1],
                                   whiten=False)
    assert list(count) == [1]
    cc = np.fft.irfft(out[0], n=n)
    a, b = x[0].astype(np.float64), x[1].astype(np.float64)
    brute = np.array([np.sum(a * np.roll(b, -tau)) for tau in range(n)])   # sum_t a[t] b[t+tau]
    assert np.allclose(cc, brute, atol=1e-9)


def test_whitened_products_are_phases_and_average():
    rng = np.random.default_rng(1)
    nsta, npts, fftlen, step = 3, 900, 128, 64
    x = rng.normal(size=(nsta, npts)).astype(np.float32)
    se = np.array([0, npts, 0, npts, 300, npts], np.int32)
    pairs = [0, 1, 0, 2, 1, 2]
    out, count = C.cross_correlate(x.reshape(-1), nsta, npts, 20, fftlen, step, pairs, se,
                                   fstride=2, whiten=True)
    nshifts = 1 + (npts - fftlen) // step
    ok2 = [k for k in range(nshifts) if k * step >= 300 and k * step + fftlen < npts]
    ok0 = [k for k in range(nshifts) if k * step + fftlen < npts]
    assert list(count) == [len(ok0), len(ok2), len(ok2)]
    # pair (0, 1): mean of unit phasors over the windows
    acc = 0
    for k in ok0:
        s = np.fft.rfft(x[:, k * step:k * step + fftlen].astype(np.float64), axis=1)[:, ::2][:, :20]
        s = s / np.abs(s)
        acc = acc + np.conj(s[0]) * s[1]
    assert np.allclose(out[0], acc / len(ok0), atol=1e-6)
    assert np.all(np.abs(out) <= 1 + 1e-6)


def test_onebit_uses_signs():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 300)).astype(np.float32)
    a, _ = C.cross_correlate(x.reshape(-1), 2, 300, 10, 128, 64, [0, 1], [0, 400, 0, 400],
                             onebit=True, whiten=False)
    b, _ = C.cross_correlate(np.sign(x).reshape(-1), 2, 300, 10, 128, 64, [0, 1],
                             [0, 400, 0, 400], whiten=False)
    assert np.array_equal(a, b)


def test_bin_stride_fold_identity():
    """Bins 0, s, 2s, ... of a length-N rFFT are the rFFT of the window
    folded to N/s samples (the device path's shortcut when s divides N)."""
    rng = np.random.default_rng(4)
    for n, s in ((14400, 3), (1024, 4), (90, 2)):
        x = rng.normal(size=n)
        full = np.fft.rfft(x)[::s]
        folded = np.fft.rfft(x.reshape(s, n // s).sum(axis=0))
        m = min(len(full), len(folded))
        assert np.allclose(full[:m], folded[:m], rtol=0, atol=1e-9 * np.abs(full).max())


def test_sign_matches_reference_rule():
    """The reference's sign() (src/CrossCorr.cpp:22-30): 1 above zero, -1
    below, 0 otherwise -- NaN fails both comparisons and gives 0, where
    np.sign would keep NaN; signed zeros give 0."""
    x = np.array([2.5, -0.1, 0.0, -0.0, np.nan, np.inf, -np.inf, 1e-45], np.float32)
    want = np.array([1, -1, 0, 0, 0, 1, -1, 1], np.float32)
    got = C.sign_ref(x)
    assert got.dtype == np.float32 and np.array_equal(got, want)
    # one-bit data with a NaN sample: the oracle's spectra stay finite
    rng = np.random.default_rng(5)
    y = rng.normal(size=(2, 300)).astype(np.float32)
    y[0, 17] = np.nan
    a, _ = C.cross_correlate(y.reshape(-1), 2, 300, 10, 128, 64, [0, 1], [0, 400, 0, 400],
                             onebit=True, whiten=False)
    assert np.all(np.isfinite(a))


def test_threaded_restatement_equals_serial():
    """The threaded CPU baseline computes what the serial restatement does,
    with uneven station spans, one-bit and a bin stride."""
    rng = np.random.default_rng(4)
    nsta, npts, fftlen, steplen, nf = 9, 700, 128, 40, 30
    x = rng.normal(size=(nsta, npts)).astype(np.float32).reshape(-1)
    se = np.tile(np.array([0, npts], np.int32), nsta)
    se[2 * 2] = 150
    se[2 * 5 + 1] = 420
    iu = np.triu_indices(nsta, 1)
    pairs = np.stack(iu, 1).astype(np.int32).reshape(-1)
    for onebit in (False, True):
        a, ca = C.cross_correlate(x, nsta, npts, nf, fftlen, steplen, pairs, se, fstride=2,
                                  onebit=onebit)
        for n in (1, 3, 8):
            b, cb = C.cross_correlate_threaded(x, nsta, npts, nf, fftlen, steplen, pairs, se,
                                               fstride=2, onebit=onebit, nthreads=n)
            assert np.array_equal(ca, cb)
            np.testing.assert_allclose(b, a, rtol=0, atol=1e-12)
