"""Python access to the CPU checkers under oracle/.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and only as the checker or the timed CPU baseline.  The
product (cc-fjpy_amd/ccfj) never imports it.

  port_fj / port_fh      FJcpu restatement (fjcpu_port.c), bit-exact with
                         the reference's FJ / FH (include/FJcpu.h:19-20)
  port_columns           the same on a subset of frequency columns
  oracle64               FP64 quadrature oracle (fj_oracle64.c)
  fj_noise, fj_earthquake, MWFJ, win
                         restatement of the reference's Python layer
                         (src/ccfj.pyx:148-268) on top of a chosen native
                         engine ("port" or "fp64"), used to check the
                         ccfj package end to end.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
REF_LIB = os.path.join(HERE, "_ref", "libfjcpu_ref.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_libs = {}


def build():
    """Compile the port and the FP64 oracle (and the reference build when
    the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.exists("/root/reference/src/FJcpu.cpp"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _lib(name):
    if name not in _libs:
        path = os.path.join(LIBDIR, name)
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        if name == "libfjport.so":
            for fn in ("fjport_FJ", "fjport_FH"):
                getattr(lib, fn).argtypes = [_P] * 5 + [_I] * 5
            lib.fjport_columns.argtypes = [_P] * 5 + [_I] * 6 + [_P, _I]
        else:
            lib.oracle64_fj.argtypes = [_P] * 4 + [_I] * 5 + [_P, _I, _P, _I]
            lib.oracle64_segment.argtypes = [ctypes.c_double] * 5 + [_I]
            lib.oracle64_segment.restype = ctypes.c_double
            lib.oracle64_j0.argtypes = [ctypes.c_double]
            lib.oracle64_j0.restype = ctypes.c_double
            lib.oracle64_y0.argtypes = [ctypes.c_double]
            lib.oracle64_y0.restype = ctypes.c_double
        _libs[name] = lib
    return _libs[name]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def host_cores():
    """Cores this process may run on (its affinity mask)."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def nthreads_default():
    """OpenMP threads of the restatements: OMP_NUM_THREADS when the host sets
    it (a share of a larger machine), else every core of the affinity mask."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return max(1, min(n, host_cores())) if n > 0 else host_cores()


def port_native(U, r, f, c, itype, ybasis, nthreads=None):
    """FJcpu restatement: FJ (ybasis 0) or FH (ybasis 1) -> float32 [nc, nf]."""
    U, r, f, c = _f32(U), _f32(r), _f32(f), _f32(c)
    nr, nf = U.shape
    out = np.zeros((len(c), nf), np.float32)
    fn = _lib("libfjport.so").fjport_FH if ybasis else _lib("libfjport.so").fjport_FJ
    fn(U.ctypes.data, r.ctypes.data, f.ctypes.data, out.ctypes.data, c.ctypes.data,
       len(c), nr, nf, int(itype), nthreads or nthreads_default())
    return out


def port_columns(U, r, f, c, itype, ybasis, fidx, nthreads=None):
    """FJcpu restatement on frequency columns fidx -> float32 [nc, len(fidx)]."""
    U, r, f, c = _f32(U), _f32(r), _f32(f), _f32(c)
    fidx = np.ascontiguousarray(fidx, dtype=np.int32)
    nr, nf = U.shape
    out = np.zeros((len(c), len(fidx)), np.float32)
    _lib("libfjport.so").fjport_columns(
        U.ctypes.data, r.ctypes.data, f.ctypes.data, out.ctypes.data, c.ctypes.data,
        len(c), nr, nf, int(itype), int(ybasis), nthreads or nthreads_default(),
        fidx.ctypes.data, len(fidx))
    return out


def oracle64(U, r, f, c, itype, ybasis, fidx=None, nthreads=None):
    """FP64 oracle -> float64 [nc, nf] (or [nc, len(fidx)])."""
    U, r, f, c = _f32(U), _f32(r), _f32(f), _f32(c)
    nr, nf = U.shape
    if fidx is None:
        nfi, fptr = nf, None
    else:
        fidx = np.ascontiguousarray(fidx, dtype=np.int32)
        nfi, fptr = len(fidx), fidx.ctypes.data
    out = np.zeros((len(c), nfi), np.float64)
    rc = _lib("liboracle64.so").oracle64_fj(
        U.ctypes.data, r.ctypes.data, f.ctypes.data, c.ctypes.data, len(c), nr, nf,
        int(itype), int(ybasis), out.ctypes.data, nthreads or nthreads_default(), fptr, nfi)
    if rc != 0:
        raise ValueError("oracle64: bad itype")
    return out


def oracle64_segment(r1, r2, g1, slope, k, ybasis):
    return _lib("liboracle64.so").oracle64_segment(r1, r2, g1, slope, k, int(ybasis))


def oracle64_j0(x):
    return _lib("liboracle64.so").oracle64_j0(x)


def oracle64_y0(x):
    return _lib("liboracle64.so").oracle64_y0(x)


# ---------------------------------------------------------------------------
# restatement of the reference's Python layer over a chosen native engine

def _engine(U, r, f, c, itype, ybasis, engine):
    if engine == "port":
        return port_native(U, r, f, c, itype, ybasis).astype(np.float64)
    return oracle64(U, r, f, c, itype, ybasis)


def _stage(uf, r, c, f, fstride):
    # src/ccfj.pyx:161-174 (staging of fj / fh)
    nr, nf = len(r), len(f)
    uf = uf[:, ::fstride].reshape(nr * nf)
    return (np.array(uf, dtype=np.float32).reshape(nr, nf), np.array(r, np.float32),
            np.array(c, np.float32), np.array(f, np.float32))


def _normalize(out):
    # src/ccfj.pyx:225-226, builtin max semantics
    for i in range(out.shape[1]):
        out[:, i] = out[:, i] / max(np.abs(out[:, i]))
    return out


def fj_noise(uf, r, c, f, fstride=1, itype=1, func=0, engine="fp64", normalize=True):
    """src/ccfj.pyx:202-227.  The Hilbert transform along frequency is
    scipy.fftpack.hilbert row by row, as there.  normalize=False returns
    the combination before the per-column division (for linearity checks)."""
    from scipy import fftpack
    indx = np.argsort(r)
    r = r[indx]
    uf = uf[indx]
    U, rr, cc, ff = _stage(uf, r, c, f, fstride)
    out = _engine(U, rr, ff, cc, itype, 0, engine)
    if func == 1:
        ufi = np.zeros(np.shape(uf), dtype=np.float32)
        for i in range(len(r)):
            ufi[i, :] = np.real(fftpack.hilbert(uf[i, :]))
        Ui, _, _, _ = _stage(ufi, r, c, f, fstride)
        out = out - _engine(Ui, rr, ff, cc, itype, 1, engine)
    if not normalize:
        return out
    with np.errstate(invalid="ignore", divide="ignore"):
        return _normalize(out)


def fj_earthquake(u, r, c, f, fstride=1, itype=1, func=0, engine="fp64", normalize=True):
    """src/ccfj.pyx:229-250 (normalize=False: before the per-column
    division)."""
    indx = np.argsort(r)
    r = r[indx]
    u1 = u[indx]
    uf = np.fft.rfft(u1)[:, 0:len(f) * fstride:fstride]
    Ur, rr, cc, ff = _stage(np.real(uf), r, c, f, fstride)
    Ui, _, _, _ = _stage(np.imag(uf), r, c, f, fstride)
    if func == 0:
        outr = _engine(Ur, rr, ff, cc, itype, 0, engine)
        outi = _engine(Ui, rr, ff, cc, itype, 0, engine)
    else:
        outr = _engine(Ur, rr, ff, cc, itype, 0, engine) - _engine(Ui, rr, ff, cc, itype, 1, engine)
        outi = _engine(Ui, rr, ff, cc, itype, 0, engine) + _engine(Ur, rr, ff, cc, itype, 1, engine)
    out = np.sqrt(outr ** 2 + outi ** 2)
    if not normalize:
        return out
    with np.errstate(invalid="ignore", divide="ignore"):
        return _normalize(out)


def win(npts, Fs, T1, T2, taper=0.8):
    """src/ccfj.pyx:148-158, element by element."""
    n1 = max(0, np.floor(T1 * Fs))
    n2 = min(npts, np.floor(T2 * Fs))
    out = np.ones(npts)
    delta = np.floor((n2 - n1) * (1 - taper) / 2)
    out[0:int(n1)] = 0
    out[int(n2):int(npts)] = 0
    for i in range(int(delta)):
        out[int(n1 + i)] = np.sin(np.pi / 2 * i / delta)
        out[int(n2 - i - 1)] = np.sin(np.pi / 2 * i / delta)
    return out


def MWFJ(u, r, c, f, Fs, nwin, winl, winr, taper=0.9, fstride=1, itype=1, func=0,
         engine="fp64"):
    """src/ccfj.pyx:252-268."""
    nr = len(r)
    npts = len(u[0])
    out = np.zeros([nwin, len(c), len(f)])
    for i in range(nwin):
        u0 = np.zeros([nr, npts])
        for j in range(nr):
            u0[j, :] = u[j, :] * win(npts, Fs, winl[i, j], winr[i, j], taper)
        out[i] = fj_earthquake(u0, r, c, f, fstride, itype, func, engine)
    return out


def column_error(x, ref):
    """max and RMS over entries of |x - ref| / max_c |ref[:, f]| (per-column
    scale, the API's normalisation).  Entries where ref is not finite must be
    non-finite in x too; they are excluded from the statistics.  Returns
    (emax, erms, n_nonfinite_mismatch)."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    fin = np.isfinite(ref)
    mismatch = int(np.sum(fin != np.isfinite(x)))
    scale = np.max(np.where(fin, np.abs(ref), 0.0), axis=0)
    scale = np.where(scale > 0, scale, 1.0)
    e = np.where(fin & np.isfinite(x), np.abs(x - ref) / scale[None, :], 0.0)
    n = max(1, int(np.sum(fin)))
    return float(e.max(initial=0.0)), float(np.sqrt(np.sum(e ** 2) / n)), mismatch
