"""CPU restatement of the reference's ambient-noise cross-correlation.
TEST INFRASTRUCTURE ONLY.

Follows src/CrossCorr.cpp:33-159 (CrossCorrelation) and src/ccfj.pyx:125-146
(the CC wrapper):
  * fftlen >= npts: fftlen = npts and one window; otherwise
    nshifts = 1 + (npts - fftlen) // steplen              (CrossCorr.cpp:50-56)
  * one-bit: data = sign(data), in place                  (:76-85)
  * window k of station i is used when k*steplen >= start_i and
    k*steplen + fftlen < end_i                             (:91-98)
  * spectrum: real FFT of the window, bins j*fstride for j < nf; whitening
    divides by max(|X|, 1e-8) in float                     (:100-119)
  * pair (a, b): sum over windows where both stations are used of
    conj(X_a) X_b, divided by that count                   (:124-150)
The reference's FFT is FFTW in single precision (its third-party dependency,
absent here); this restatement uses numpy's float64 FFT of the float32 data,
so the two agree to single-precision rounding, not bit for bit.
"""
import numpy as np


def windows(npts, fftlen, steplen):
    """(fftlen, nshifts) after the reference's clamp."""
    if fftlen >= npts:
        return npts, 1
    return fftlen, 1 + (npts - fftlen) // steplen


def used(startend, k, steplen, fftlen):
    """Stations whose window k lies inside [start, end)."""
    se = np.asarray(startend).reshape(-1, 2)
    t0 = k * steplen
    return (t0 >= se[:, 0]) & (t0 + fftlen < se[:, 1])


def sign_ref(x):
    """The reference's sign() (src/CrossCorr.cpp:22-30): 1 for x > 0, -1
    for x < 0, else 0 -- a NaN sample fails both comparisons and becomes 0
    (np.sign would keep NaN)."""
    x = np.asarray(x, np.float32)
    return np.where(x > 0, np.float32(1), np.where(x < 0, np.float32(-1), np.float32(0)))


def cross_correlate(data, nsta, npts, nf, fftlen, steplen, pairs, startend, fstride=1,
                    onebit=False, whiten=True):
    """-> (ncfs complex128 [npairs, nf], count [npairs])."""
    x = np.asarray(data, np.float32).reshape(nsta, npts)
    if onebit:
        x = sign_ref(x)
    fftlen, nshifts = windows(npts, fftlen, steplen)
    p = np.asarray(pairs).reshape(-1, 2)
    acc = np.zeros((len(p), nf), np.complex128)
    count = np.zeros(len(p), np.int64)
    for k in range(nshifts):
        ok = used(startend, k, steplen, fftlen)
        seg = x[:, k * steplen:k * steplen + fftlen].astype(np.float64)
        spec = np.fft.rfft(seg, n=fftlen, axis=1)[:, ::fstride][:, :nf]
        if whiten:
            mag = np.maximum(np.abs(spec).astype(np.float32), np.float32(1e-8))
            spec = spec / mag
        both = ok[p[:, 0]] & ok[p[:, 1]]
        acc[both] += np.conj(spec[p[both, 0]]) * spec[p[both, 1]]
        count[both] += 1
    out = np.zeros_like(acc)
    has = count > 0
    out[has] = acc[has] / count[has, None]
    return out, count


def cross_correlate_threaded(data, nsta, npts, nf, fftlen, steplen, pairs, startend, fstride=1,
                             onebit=False, whiten=True, nthreads=1):
    """cross_correlate on nthreads host threads, as the reference spreads the
    station FFTs and the pair sums over OpenMP threads (CrossCorr.cpp:86-138):
    the window FFTs of all stations run in scipy's threaded real FFT and the
    pairs are summed in nthreads contiguous chunks.  Same result as
    cross_correlate up to float64 rounding."""
    import scipy.fft
    from concurrent.futures import ThreadPoolExecutor
    x = np.asarray(data, np.float32).reshape(nsta, npts)
    if onebit:
        x = sign_ref(x)
    fftlen, nshifts = windows(npts, fftlen, steplen)
    p = np.asarray(pairs).reshape(-1, 2)
    acc = np.zeros((len(p), nf), np.complex128)
    count = np.zeros(len(p), np.int64)
    bounds = np.linspace(0, len(p), nthreads + 1).astype(np.int64)

    def pair_sum(q, spec, ok):
        lo, hi = bounds[q], bounds[q + 1]
        pa, pb = p[lo:hi, 0], p[lo:hi, 1]
        both = ok[pa] & ok[pb]
        idx = np.nonzero(both)[0]
        acc[lo + idx] += np.conj(spec[pa[idx]]) * spec[pb[idx]]
        count[lo + idx] += 1

    with ThreadPoolExecutor(nthreads) as pool:
        for k in range(nshifts):
            ok = used(startend, k, steplen, fftlen)
            seg = x[:, k * steplen:k * steplen + fftlen].astype(np.float64)
            spec = scipy.fft.rfft(seg, n=fftlen, axis=1, workers=nthreads)[:, ::fstride][:, :nf]
            if whiten:
                mag = np.maximum(np.abs(spec).astype(np.float32), np.float32(1e-8))
                spec = spec / mag
            spec = np.ascontiguousarray(spec)
            list(pool.map(lambda q: pair_sum(q, spec, ok), range(nthreads)))
    out = np.zeros_like(acc)
    has = count > 0
    out[has] = acc[has] / count[has, None]
    return out, count
