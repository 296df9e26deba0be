"""Seeded synthetic inputs for the F-J workloads (no data files needed).

Each generator returns the arrays the reference's Python API takes, in the
units it uses (r in metres, c in m/s, f in Hz).  Constants the workload
descriptions leave open are fixed here.

  noise_single(seed=0)   300 CCFs, r ~ U[5,200] km, nf=512 (0.01-0.5 Hz),
                         nc=500 (1000-5000 m/s); uf = J0(2 pi f r / c0(f))
                         + complex N(0, 0.1), c0 = 3000 + 1500 exp(-f/0.08)
  noise_dense(seed=1)    200 stations in a 300 x 300 km square -> 19,900
                         pairs (GetStationPairs order, r unsorted),
                         nf=4096 (0.005-2 Hz), nc=3000 (500-6000 m/s);
                         three modes, amplitudes 1 / 0.5 / 0.25,
                         c_m(f) = cinf_m + d_m exp(-f/phi_m), plus noise.
                         Only the real part of uf is used by the transform,
                         so `real_only=True` skips the imaginary noise.
  subarrays(seed=3)      4096 sub-arrays of 20-400 pairs, nf=512 (0.02-1 Hz),
                         nc=1024 (1000-5000 m/s), two modes each.
  active_source(seed=2)  480 receivers at 1-120 km, 16384 samples at 100 Hz:
                         three surface-wave modes from a 1 Hz Ricker source
                         with 1/sqrt(r) spreading plus 5% white noise, and 64
                         group-velocity windows (equal width, 50% overlap)
                         tiling 1.5-4.5 km/s; nc=2000 (1500-4500 m/s),
                         f = k Fs/npts for k < 820 (0-5 Hz).
"""
import os

import numpy as np
from scipy.special import j0

DENSE_MODES = dict(cinf=(2000.0, 2600.0, 3200.0), delta=(2000.0, 2400.0, 2600.0),
                   phi=(0.15, 0.3, 0.5), amp=(1.0, 0.5, 0.25))


def noise_single(seed=0, nr=300, nf=512, nc=500):
    rng = np.random.default_rng(seed)
    r = np.sort(rng.uniform(5e3, 200e3, nr))
    f = np.linspace(0.01, 0.5, nf)
    c = np.linspace(1000.0, 5000.0, nc)
    c0 = 3000.0 + 1500.0 * np.exp(-f / 0.08)
    uf = j0(2 * np.pi * f[None, :] * r[:, None] / c0[None, :])
    uf = uf + rng.normal(0, 0.1, uf.shape) + 1j * rng.normal(0, 0.1, uf.shape)
    return dict(uf=uf.astype(np.complex64), r=r, c=c, f=f, c0=c0)


def dense_pairs(seed=1, nsta=200, side=300e3):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0.0, side, (nsta, 2))
    ii, jj = np.triu_indices(nsta, k=1)      # GetStationPairs order (i<j, row-major)
    r = np.hypot(*(xy[ii] - xy[jj]).T)
    return r, rng


def dense_phase_velocity(f):
    m = DENSE_MODES
    return [ci + d * np.exp(-f / p) for ci, d, p in zip(m["cinf"], m["delta"], m["phi"])]


def noise_dense(seed=1, nf=4096, nc=3000, real_only=False, chunk=512, fidx=None):
    """fidx (optional): keep only these frequency columns (the noise is
    drawn for the full grid first, so a column is identical either way)."""
    r, rng = dense_pairs(seed)
    f = np.linspace(0.005, 2.0, nf)
    c = np.linspace(500.0, 6000.0, nc)
    if fidx is not None:
        full = rng.normal(0, 0.1, (len(r), nf)).astype(np.float32)[:, fidx]
        f = f[fidx]
        nf = len(f)
    cm = dense_phase_velocity(f)
    nr = len(r)
    re = np.empty((nr, nf), np.float32)
    for a in range(0, nr, chunk):
        rr = r[a:a + chunk, None]
        acc = np.zeros((len(rr), nf))
        for amp, cv in zip(DENSE_MODES["amp"], cm):
            acc += amp * j0(2 * np.pi * f[None, :] * rr / cv[None, :])
        re[a:a + chunk] = acc
    if fidx is not None:
        re += full
        return dict(uf=re, r=r, c=c, f=f, cm=cm)
    re += rng.normal(0, 0.1, re.shape).astype(np.float32)
    if real_only:
        uf = re
    else:
        uf = re + 1j * rng.normal(0, 0.1, re.shape).astype(np.float32)
        uf = uf.astype(np.complex64)
    return dict(uf=uf, r=r, c=c, f=f, cm=cm)


def subarrays(seed=3, nb=4096, nmin=20, nmax=400, nf=512, nc=1024, real_only=True):
    """Many independent sub-arrays: nr_b ~ U{nmin..nmax}, r sorted in
    [1, 150] km, f 0.02-1 Hz, c 1000-5000 m/s; two random monotone
    decreasing dispersion curves per sub-array (cinf in [1800, 3600] m/s,
    rise in [500, 2500] m/s, decay in [0.1, 0.6] Hz, amplitudes 1 and
    U[0.3, 0.7]) plus noise sigma 0.1.  Returns lists of per-sub-array
    spectra and distances, and the shared grids."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    f = np.linspace(0.02, 1.0, nf)
    c = np.linspace(1000.0, 5000.0, nc)
    ns = rng.integers(nmin, nmax + 1, nb)
    # the random draws stay in one sequence (sub-array by sub-array); the
    # Bessel sums, which release the GIL, run on worker threads
    jobs, rs = [], []
    for n in ns:
        r = np.sort(rng.uniform(1e3, 150e3, n))
        modes = []
        for amp in (1.0, rng.uniform(0.3, 0.7)):
            cv = rng.uniform(1800, 3600) + rng.uniform(500, 2500) * np.exp(-f / rng.uniform(0.1, 0.6))
            modes.append((amp, cv))
        noise = rng.normal(0, 0.1, (n, nf))
        inoise = None if real_only else rng.normal(0, 0.1, (n, nf))
        jobs.append((r, modes, noise, inoise))
        rs.append(r)

    def one(job):
        r, modes, noise, inoise = job
        acc = np.zeros((len(r), nf))
        for amp, cv in modes:
            acc += amp * j0(2 * np.pi * f[None, :] * r[:, None] / cv[None, :])
        acc += noise
        if inoise is not None:
            acc = acc + 1j * inoise
        return acc.astype(np.float32 if inoise is None else np.complex64)

    try:
        workers = max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        workers = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(workers) as ex:
        ufs = list(ex.map(one, jobs))
    return dict(ufs=ufs, rs=rs, c=c, f=f, ns=ns)


ACTIVE_MODES = dict(cinf=(2200.0, 3000.0, 3600.0), delta=(1400.0, 1300.0, 800.0),
                    phi=(0.5, 1.0, 1.5), amp=(1.0, 0.6, 0.3))


def active_source(seed=2, nr=480, npts=16384, Fs=100.0, nwin=64, nf=820, nc=2000):
    """Multi-window workload: u float32 [nr, npts], r [nr] (m), winl/winr
    [nwin, nr] (s), and the grids f, c, Fs."""
    rng = np.random.default_rng(seed)
    r = np.linspace(1e3, 120e3, nr)
    fa = np.fft.rfftfreq(npts, 1.0 / Fs)
    fp = 1.0
    ricker = (2.0 / np.sqrt(np.pi)) * fa ** 2 / fp ** 3 * np.exp(-(fa / fp) ** 2)
    spec = np.zeros((nr, len(fa)), np.complex128)
    m = ACTIVE_MODES
    for ci, d, ph, a in zip(m["cinf"], m["delta"], m["phi"], m["amp"]):
        cm = ci + d * np.exp(-fa / ph)
        spec += a * ricker[None, :] * np.exp(-2j * np.pi * fa[None, :] * r[:, None] / cm[None, :])
    u = np.fft.irfft(spec, n=npts, axis=1) / np.sqrt(r)[:, None]
    rms = np.sqrt(np.mean(u ** 2))
    u += rng.normal(0.0, 0.05 * rms, u.shape)
    w = 3.0 / (nwin + 1) * 2.0          # window width, km/s; steps of w/2 tile [1.5, 4.5]
    vlo = 1.5 + np.arange(nwin) * w / 2
    vhi = vlo + w
    winl = r[None, :] / (vhi[:, None] * 1e3)
    winr = r[None, :] / (vlo[:, None] * 1e3)
    f = np.arange(nf) * Fs / npts
    c = np.linspace(1500.0, 4500.0, nc)
    return dict(u=u.astype(np.float32), r=r, winl=winl, winr=winr, f=f, c=c, Fs=Fs,
                nwin=nwin)


def term_evals(nr, nf, nc, calls=1):
    """Reference-equivalent term evaluations: nf * nc * nr per native call."""
    return float(nr) * nf * nc * calls
