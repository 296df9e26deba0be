"""ccfj -- MI355X-native frequency-Bessel (F-J) transform, drop-in for CC-FJpy.

The public functions keep the names, defaults, argument meaning, dtypes and
output layout of the reference's Python API (src/ccfj-gpu.pyx:168-313, the
GPU build; src/ccfj.pyx:148-268, the CPU build):

    fj_noise(uf, r, c, f, fstride=1, itype=1, func=0, num=-1)   -> float32 [nc, nf]
    fj_earthquake(u, r, c, f, fstride=1, itype=1, func=0, num=-1) -> float32 [nc, nf]
    MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, fstride=1,
         itype=1, func=0, num=-1)                                -> float64 [nwin, nc, nf]
    fj, fh, fhr, fhi (un-normalised single-basis transforms), win,
    GetStationPairs
    CC(npts, nsta, nf, fftlen, Pairs, startend, data, overlaprate=0.0,
       nThreads=8, fstride=1, ifonebit=0, ifspecwhittenning=1)  -> complex64 [npairs, nf]

`num` selects the GPU (-1 = the current device); the reference's GPU build
did this through CUDA_VISIBLE_DEVICES (src/ccfj-gpu.pyx:247-248).

All F-J arithmetic runs in the HIP kernels of libccfj.so; this module only
sorts, casts, slices and combines as the reference's Python layer does.
There is no CPU fallback: importing the package without its native parts,
or calling it without a GPU, raises.
"""
import os
import warnings

import numpy as np

try:
    from . import _ccfj
except ImportError as exc:  # the native extension is the product; no fallback
    raise ImportError(
        "ccfj: native extension not built (run `python cc-fjpy_amd/build.py`): %s" % exc
    ) from exc

__all__ = ["fj_noise", "fj_earthquake", "MWFJ", "fj", "fh", "fhr", "fhi", "win",
           "GetStationPairs", "CC", "hilbert_rows", "device_count", "fj_noise_batch",
           "fj_full", "fH_full", "fHr_full", "fHi_full", "CC_full"]

MODE_J = _ccfj.MODE_J
MODE_Y = _ccfj.MODE_Y
MODE_J_MINUS_Y = _ccfj.MODE_J_MINUS_Y
MODE_ABS_J = _ccfj.MODE_ABS_J
MODE_ABS_H = _ccfj.MODE_ABS_H

_contexts = {}


def device_count():
    return _ccfj.device_count()


def _context(num=-1):
    """Per-device context (stream + cached device workspace)."""
    dev = -1 if num is None or int(num) < 0 else int(num)
    ctx = _contexts.get(dev)
    if ctx is None:
        ctx = _ccfj.Context(dev)
        _contexts[dev] = ctx
    return ctx


def _as_f32(a):
    # the reference casts with np.array(.., dtype=np.float32) only when the
    # dtype differs (src/ccfj-gpu.pyx:188-195); complex input keeps its real
    # part and numpy emits ComplexWarning, exactly as there
    if a.dtype != np.float32:
        a = np.array(a, dtype=np.float32)
    return np.ascontiguousarray(a)


def _stage(uf, r, c, f, fstride):
    """Staging of the reference's fj/fh (src/ccfj-gpu.pyx:181-197)."""
    nr = len(r)
    nf = len(f)
    uf = uf[:, ::fstride]
    uf = uf.reshape(nr * nf)
    r = _as_f32(np.asarray(r))
    f = _as_f32(np.asarray(f))
    uf = _as_f32(uf).reshape(nr, nf)
    c = _as_f32(np.asarray(c))
    return uf, r, c, f


def _native(mode, itype, normalize, U0, U1, r, c, f, num):
    nr, nf = len(r), len(f)
    if itype not in (0, 1):
        # the reference's native code silently does nothing for other
        # itype values (src/FJcpu.cpp:403-407); the output stays zero
        out = np.zeros((len(c), nf), np.float32)
        if normalize:
            with np.errstate(invalid="ignore", divide="ignore"):
                _normalize_columns(out)
        return out
    ctx = _context(num)
    return ctx.transform(mode, int(itype), bool(normalize), U0.reshape(nr, nf),
                         None if U1 is None else U1.reshape(nr, nf), r, f, c)


def _normalize_columns(out):
    # src/ccfj.pyx:225-226; Python's builtin max over the column
    for i in range(out.shape[1]):
        out[:, i] = out[:, i] / max(np.abs(out[:, i]))


# ---- the reference's module-level native shims (typed float32 / int32
# memoryviews in, results written into `out` / ncfsr, ncfsi; a native error
# raises RuntimeError).  They run on the current device.

def fj_full(U_f, r, f, out, c, nc, nr, nf, type, nThread=20):
    """J0 basis into out[nc * nf] (src/ccfj.pyx:72-90 CPU form with
    nThread, accepted and ignored; src/ccfj-gpu.pyx:75-92 GPU form)."""
    _ccfj.fj_full(U_f, r, f, out, c, nc, nr, nf, type)


def fH_full(U_f, r, f, out, c, nc, nr, nf, type, nThread=20):
    """Y0 basis into out[nc * nf] (src/ccfj.pyx:95-113; nThread ignored)."""
    _ccfj.fH_full(U_f, r, f, out, c, nc, nr, nf, type, nThread)


def fHr_full(U_f, r, f, out, c, nc, nr, nf, type):
    """J0 basis into out[nc * nf] (src/ccfj-gpu.pyx:97-114)."""
    _ccfj.fHr_full(U_f, r, f, out, c, nc, nr, nf, type)


def fHi_full(U_f, r, f, out, c, nc, nr, nf, type):
    """Y0 basis into out[nc * nf] (src/ccfj-gpu.pyx:118-135)."""
    _ccfj.fHi_full(U_f, r, f, out, c, nc, nr, nf, type)


def CC_full(npts, nsta, nf, fstride, fftlen, steplen, ifonebit, ifspecwhitenning, nThreads,
            npairs, Pairs, startend, data, ncfsr, ncfsi):
    """Cross-correlation into ncfsr/ncfsi[npairs * nf], added to what they
    hold (src/ccfj.pyx:44-68); nThreads ignored."""
    _ccfj.CC_full(npts, nsta, nf, fstride, fftlen, steplen, ifonebit, ifspecwhitenning,
                  nThreads, npairs, Pairs, startend, data, ncfsr, ncfsi)


def fj(uf, r, c, f, fstride=1, itype=1, nThread=20, num=-1):
    """J0-basis transform of uf (no normalisation), float32 [nc, nf]
    (src/ccfj-gpu.pyx:181-197; the CPU build's nThread, src/ccfj.pyx:161,
    is accepted and ignored)."""
    U, r, c, f = _stage(uf, r, c, f, fstride)
    return _native(MODE_J, itype, False, U, None, r, c, f, num)


def fhr(uf, r, c, f, fstride=1, itype=1, num=-1):
    """J0-basis transform (the GPU build's FP64 "real Hankel" part)."""
    return fj(uf, r, c, f, fstride, itype, num=num)


def fhi(uf, r, c, f, fstride=1, itype=1, num=-1):
    """Y0-basis transform of uf (no normalisation), float32 [nc, nf]."""
    U, r, c, f = _stage(uf, r, c, f, fstride)
    return _native(MODE_Y, itype, False, U, None, r, c, f, num)


def fh(uf, r, c, f, fstride=1, itype=1, nThread=20, num=-1):
    """Y0-basis transform under the CPU build's name (src/ccfj.pyx:181-198)."""
    return fhi(uf, r, c, f, fstride, itype, num)


def hilbert_rows(x):
    """Row-wise Hilbert transform with scipy.fftpack's convention, as the
    reference applies it along frequency (src/ccfj.pyx:217-219):
    y_k = i sign(k) x_k in Fourier space, with the zero and (even length)
    Nyquist coefficients set to zero, so hilbert(cos) = -sin.  A complex
    row is transformed part by part; the float32 store then keeps
    hilbert(Re x)."""
    x = np.asarray(x)
    if np.iscomplexobj(x):
        x = x.real
    n = x.shape[-1]
    X = np.fft.fft(x.astype(np.float64, copy=False), axis=-1)
    k = np.fft.fftfreq(n) * n
    h = 1j * np.sign(k)
    if n % 2 == 0:
        h[n // 2] = 0.0
    return np.fft.ifft(X * h, axis=-1).real.astype(np.float32)


def fj_noise(uf, r, c, f, fstride=1, itype=1, func=0, num=-1):
    """F-J of noise cross-correlation spectra (src/ccfj-gpu.pyx:241-268).

    uf: [nr, ncol] spectra (only the real part is used), r: distances (m),
    c: phase velocities (m/s), f: frequencies, ceil(ncol/fstride) == len(f).
    func 0: Bessel basis; func 1: Hankel basis, FJ[uf] - FY[Hilbert_f(uf)]
    with the Hilbert transform taken along frequency on the GPU.
    Returns the per-frequency normalised float32 [nc, nf] spectrum.
    """
    indx = np.argsort(r)
    r = r[indx]
    fast = (isinstance(uf, np.ndarray) and uf.ndim == 2 and uf.flags.c_contiguous
            and uf.dtype in (np.complex64, np.float32))
    if not fast:
        uf = uf[indx]
    if func not in (0, 1):
        print('set func as 0 for Bessel function 1 for Hankel function')
        return 0
    nr, nf = len(r), len(f)
    uf[:, ::fstride].reshape(nr * nf)      # the reference's shape check (ValueError)
    if itype not in (0, 1):
        return _native(MODE_J, itype, True, None, None, r, c, f, num)
    if fast:
        # the row sort and the float32 cast in one threaded pass; numpy's
        # cast warning for complex input, as the reference's cast gives it
        if uf.dtype == np.complex64:
            warnings.warn("Casting complex values to real discards the imaginary part",
                          np.exceptions.ComplexWarning, stacklevel=2)
        U = _ccfj.gather_real_rows(uf, indx)
    else:
        U = _as_f32(np.asarray(uf))
    return _context(num).fj_noise(int(itype), int(func), int(fstride), U.reshape(nr, -1),
                                  _as_f32(np.asarray(r)), _as_f32(np.asarray(f)),
                                  _as_f32(np.asarray(c)))


def _earthquake_spectra(u, r, f, fstride):
    indx = np.argsort(r)
    r = r[indx]
    u1 = u[indx]
    uf = np.fft.rfft(u1)
    uf = uf[:, 0:len(f) * fstride:fstride]
    return np.real(uf), np.imag(uf), r


def fj_earthquake(u, r, c, f, fstride=1, itype=1, func=0, num=-1):
    """F-J of earthquake / active-source records (src/ccfj-gpu.pyx:270-295).

    u: [nr, npts] time series; the rFFT along time is sliced to
    0:len(f)*fstride:fstride and then staged like fj (which strides again,
    as the reference does).  func 0: |FJ[Re] + i FJ[Im]|; func 1:
    |FJ[Re] - FY[Im] + i (FJ[Im] + FY[Re])|.  Normalised float32 [nc, nf].
    """
    ufr, ufi, r = _earthquake_spectra(u, r, f, fstride)
    if func not in (0, 1):
        print('set func as 0 for Bessel function 1 for Hankel function')
        return 0
    Ur, rr, cc, ff = _stage(ufr, r, c, f, fstride)
    Ui, _, _, _ = _stage(ufi, r, c, f, fstride)
    mode = MODE_ABS_J if func == 0 else MODE_ABS_H
    return _native(mode, itype, True, Ur, Ui, rr, cc, ff, num)


def win(npts, Fs, T1, T2, taper=0.8):
    """Tapered box window (src/ccfj.pyx:148-158): zero outside
    [floor(T1 Fs), floor(T2 Fs)), sine ramps of floor((n2-n1)(1-taper)/2)
    samples at both ends."""
    n1 = max(0, np.floor(T1 * Fs))
    n2 = min(npts, np.floor(T2 * Fs))
    delta = np.floor((n2 - n1) * (1 - taper) / 2)
    i1, i2, d = int(n1), int(n2), int(delta)
    if 0 <= i1 <= i2 <= npts and 2 * d <= i2 - i1:
        out = np.ones(npts)
        out[0:i1] = 0
        out[i2:int(npts)] = 0
        if d > 0:
            # np.pi/2*i/delta in the reference's scalar arithmetic: a numpy
            # float delta (float32 window edges give one) sets the precision
            dt = delta.dtype if isinstance(delta, np.floating) else np.dtype(np.float64)
            ramp = np.sin((np.pi / 2 * np.arange(d)).astype(dt) / dt.type(delta))
            out[i1:i1 + d] = ramp
            out[i2 - d:i2] = ramp[::-1]
        return out
    # degenerate windows (reversed, out of range, overlapping ramps): the
    # element-by-element semantics of the reference, negative indices included
    out = np.ones(npts)
    out[0:i1] = 0
    out[i2:int(npts)] = 0
    for i in range(d):
        out[int(n1 + i)] = np.sin(np.pi / 2 * i / delta)
        out[int(n2 - i - 1)] = np.sin(np.pi / 2 * i / delta)
    return out


def _window_params(T1, T2, npts, Fs, taper):
    """(n1, n2, delta) of win() for arrays of window edges, in the
    arithmetic of the reference's scalar expressions (src/ccfj.pyx:149-152):
    numpy promotes them as it promotes the arrays, so float32 edges give
    float32 products, differences and floors (which can differ from float64
    by one sample); max(0, x) and min(npts, x) keep the Python int unless x
    is larger / smaller, and with both ends clamped delta is computed from
    those ints, in double."""
    p1 = np.floor(T1 * Fs)
    p2 = np.floor(T2 * Fs)
    c1 = ~(p1 > 0)
    c2 = ~(p2 < npts)
    n1 = np.where(c1, np.zeros((), p1.dtype), p1)
    n2 = np.where(c2, np.asarray(npts, p2.dtype), p2)
    d = np.floor((n2 - n1) * (1 - taper) / 2)
    both = c1 & c2
    if np.any(both):
        d = np.where(both, np.floor((npts - 0) * (1 - taper) / 2), d)
    return n1, n2, d


def _window_descriptions(winl, winr, indx, npts, Fs, taper):
    """Integer form of win() for every (window, sorted station): zero
    outside [n1, n2), sine ramps of delta samples.  Windows that win()
    builds through its degenerate branches (reversed, out of range,
    overlapping ramps) are returned as explicit rows instead."""
    T1 = np.asarray(winl)[:, indx]
    T2 = np.asarray(winr)[:, indx]
    a, b, d = _window_params(T1, T2, npts, Fs, taper)
    regular = (a >= 0) & (a <= b) & (b <= npts) & (2 * d <= b - a)
    n1 = np.ascontiguousarray(a, dtype=np.int32)
    n2 = np.ascontiguousarray(b, dtype=np.int32)
    dl = np.ascontiguousarray(np.where(regular, d, 0), dtype=np.int32)
    xidx = np.full(T1.shape, -1, np.int32)
    rows = []
    for w, s in zip(*np.nonzero(~regular)):
        xidx[w, s] = len(rows)
        rows.append(win(npts, Fs, T1[w, s], T2[w, s], taper))
    xrows = np.array(rows, np.float64).reshape(len(rows), npts)
    return n1, n2, dl, xidx, xrows


def MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, fstride=1, itype=1, func=0,
         num=-1):
    """Multi-window F-J (src/ccfj-gpu.pyx:297-313): for each window w the
    traces are tapered by win(npts, Fs, winl[w,j], winr[w,j], taper) (window
    edges in seconds, original station order) and transformed as
    fj_earthquake does.  Tapering, the rFFT and the F-J of all windows run on
    the GPU (itype 0: windows as right-hand sides of one MFMA contraction per
    frequency).  Returns float64 [nwin, nc, nf]."""
    nr = len(r)
    npts = len(u[0])
    nf, nc = len(f), len(c)
    if func not in (0, 1):
        for _ in range(nwin):
            print('set func as 0 for Bessel function 1 for Hankel function')
        return np.zeros([nwin, nc, nf])
    # fj_earthquake slices bins 0:nf*fstride:fstride and fj slices again
    # before reshape(nr*nf) (src/ccfj.pyx:234, :165-166)
    ncols = len(range(0, min(nf * fstride, npts // 2 + 1), fstride))
    ncols2 = len(range(0, ncols, fstride))
    if ncols2 != nf:
        raise ValueError("cannot reshape array of size %d into shape (%d,)" % (nr * ncols2, nr * nf))
    if itype not in (0, 1):
        return np.full([nwin, nc, nf], np.nan)
    indx = np.argsort(r)
    rs = np.asarray(r)[indx]
    us = np.ascontiguousarray(np.asarray(u)[indx], dtype=np.float64)
    n1, n2, dl, xidx, xrows = _window_descriptions(np.asarray(winl)[:nwin], np.asarray(winr)[:nwin],
                                                   indx, npts, Fs, taper)
    return _context(num).mwfj(int(itype), int(func), int(fstride), us, _as_f32(rs),
                              _as_f32(np.asarray(f)), _as_f32(np.asarray(c)), n1, n2, dl, xidx,
                              xrows)


def fj_noise_batch(ufs, rs, c, f, fstride=1, itype=1, func=0, num=-1):
    """fj_noise over many independent sub-arrays in one persistent launch.

    No reference counterpart (the reference transforms one array per call).
    ufs, rs: sequences of per-sub-array spectra [nr_b, ncol] and distances
    [nr_b]; all sub-arrays share c and f.  Each sub-array is sorted, staged
    and combined exactly as fj_noise does.  Returns float32 [nb, nc, nf],
    each sub-array normalised per frequency column.
    """
    if func not in (0, 1):
        print('set func as 0 for Bessel function 1 for Hankel function')
        return 0
    nb = len(ufs)
    if len(rs) != nb:
        raise ValueError("ufs and rs differ in length")
    Us, Hs, Rs = [], [], []
    cc = ff = None
    for uf, r in zip(ufs, rs):
        r = np.asarray(r)
        indx = np.argsort(r)
        r = r[indx]
        uf = np.asarray(uf)[indx]
        U, rr, cc, ff = _stage(uf, r, c, f, fstride)
        Us.append(U)
        Rs.append(rr)
        if func == 1:
            Hs.append(_stage(hilbert_rows(uf), r, c, f, fstride)[0])
    soff = np.zeros(nb + 1, np.int64)
    soff[1:] = np.cumsum([len(r) for r in Rs])
    if itype not in (0, 1):
        out = np.zeros((nb, len(cc), len(ff)), np.float32)
        with np.errstate(invalid="ignore", divide="ignore"):
            for b in range(nb):
                _normalize_columns(out[b])
        return out
    U = np.ascontiguousarray(np.concatenate(Us, axis=0))
    H = np.ascontiguousarray(np.concatenate(Hs, axis=0)) if func == 1 else None
    R = np.ascontiguousarray(np.concatenate(Rs))
    mode = MODE_J if func == 0 else MODE_J_MINUS_Y
    return _context(num).transform_batched(mode, int(itype), True, soff, U, H, R, ff, cc)


def CC(npts, nsta, nf, fftlen, Pairs, startend, data, overlaprate=0.0, nThreads=8, fstride=1,
       ifonebit=0, ifspecwhittenning=1, num=-1):
    """Ambient-noise cross-correlation (src/ccfj.pyx:125-146 with
    src/CrossCorr.cpp:33-159), on the GPU: windows of fftlen samples every
    int(fftlen (1 - overlaprate)) samples, one-bit and spectral whitening
    as flagged, the mean over shared windows of conj(X_a) X_b at bins
    0, fstride, ... (nf of them) for every pair of Pairs (flattened, as
    GetStationPairs returns it).  data: the nsta records of npts samples,
    flattened; startend[2i:2i+2]: the samples of station i that hold data.
    As the reference, a float32 `data` is replaced by its signs in place when
    ifonebit is set.  nThreads is accepted and ignored.  Returns complex64
    [npairs, nf]."""
    steplen = int(fftlen * (1 - overlaprate))
    npairs = int(len(Pairs) / 2)
    if data.dtype != np.float32:
        data = np.array(data, dtype=np.float32)   # the reference's copy
    out = _context(num).cross_correlate_c64(int(nsta), int(npts), int(nf), int(fstride),
                                            int(fftlen), steplen, bool(ifonebit),
                                            bool(ifspecwhittenning),
                                            np.asarray(Pairs, np.int32)[:2 * npairs],
                                            np.asarray(startend, np.int32), data)
    if ifonebit:
        # the reference's sign() (src/CrossCorr.cpp:22-30): a NaN sample
        # fails both comparisons and becomes 0
        with np.errstate(invalid="ignore"):
            np.copyto(data, np.where(data > 0, np.float32(1),
                                     np.where(data < 0, np.float32(-1), np.float32(0))))
    return out


def GetStationPairs(nsta):
    """Station pair list i<j, flattened (src/ccfj.pyx:116-123)."""
    pairs = [(ii, jj) for ii in range(nsta) for jj in range(ii + 1, nsta)]
    return np.array(pairs, dtype=np.int32).reshape(-1)
