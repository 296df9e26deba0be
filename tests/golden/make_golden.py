"""Write the FJcpu golden vectors (tests/golden/fjcpu_golden.npz).

Run in a container that holds the reference tree:

    make -C oracle ref            # compiles src/FJcpu.cpp alone into oracle/_ref/
    python tests/golden/make_golden.py

The reference's native F-J entry points FJ/FH (include/FJcpu.h:19-20) are
called through ctypes on small, edge-case-rich float32 inputs; inputs and
outputs are stored as plain arrays.  The reference source itself is never
copied: only the numbers it produced.

Cases (each stores U [nr,nf], r [nr], f [nf], c [nc] and the four native
outputs fj0, fj1, fh0, fh1 = FJ/FH with type 0/1, each [nc,nf]):
  nr2        two stations
  nr3        three stations, one gap below the 0.1 m clamp
  edges17    17 stations: a duplicated r, two gaps < 0.1 m, non-uniform
             spacing, and an f = 0 column (NaN for type 1 in the reference)
  branches   40 stations from 10 m to 120 km so that x = k r spans the
             x<8, 8-20, 20-50 and >50 branches of the special functions
  cfg1s      the 300-pair noise field of the first benchmark workload
             (J0 dispersion model + noise, complex64 reduced to its real part
             by the float32 cast) on 16 frequencies x 24 velocities
nf >= 10 everywhere: the reference's trap_J prints f[0..9].
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libfjcpu_ref.so")


def _call(lib, name, U, r, f, c, itype):
    nr, nf = U.shape
    nc = len(c)
    out = np.zeros(nc * nf, np.float32)
    fn = getattr(lib, name)
    fn.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int] * 5
    fn.restype = ctypes.c_int
    U = np.ascontiguousarray(U, np.float32)
    fn(U.ctypes.data, r.ctypes.data, f.ctypes.data, out.ctypes.data,
       c.ctypes.data, nc, nr, nf, itype, 4)
    return out.reshape(nc, nf)


def cases():
    rng = np.random.default_rng(20261203)
    out = {}

    r = np.array([1500.0, 4200.0], np.float32)
    f = np.linspace(0.05, 1.0, 10).astype(np.float32)
    c = np.linspace(800, 4000, 8).astype(np.float32)
    U = rng.standard_normal((2, 10)).astype(np.float32)
    out["nr2"] = (U, r, f, c)

    r = np.array([2000.0, 2000.05, 9000.0], np.float32)
    f = np.linspace(0.02, 0.6, 11).astype(np.float32)
    c = np.linspace(1000, 5000, 9).astype(np.float32)
    U = rng.standard_normal((3, 11)).astype(np.float32)
    out["nr3"] = (U, r, f, c)

    r = np.sort(rng.uniform(300.0, 60e3, 17)).astype(np.float64)
    r[6] = r[5]                 # duplicate distance
    r[10] = r[9] + 0.04         # gap below the 0.1 m clamp
    r[13] = r[12] + 0.09
    r = np.sort(r).astype(np.float32)
    f = np.concatenate([[0.0], np.geomspace(0.01, 1.5, 11)]).astype(np.float32)
    c = np.geomspace(300, 6000, 10).astype(np.float32)
    U = rng.standard_normal((17, 12)).astype(np.float32)
    out["edges17"] = (U, r, f, c)

    r = np.sort(np.geomspace(10.0, 120e3, 40)).astype(np.float32)
    f = np.geomspace(0.001, 2.0, 14).astype(np.float32)
    c = np.geomspace(300, 6000, 12).astype(np.float32)
    U = (rng.standard_normal((40, 14)) * 0.3 + 1.0).astype(np.float32)
    out["branches"] = (U, r, f, c)

    # first benchmark workload: 300 pairs, r in [5, 200] km, J0 dispersion
    # model c0(f) = 3000 + 1500 exp(-f/0.08) plus complex noise sigma 0.1
    from scipy.special import j0
    rr = np.sort(rng.uniform(5e3, 200e3, 300))
    ff = np.linspace(0.01, 0.5, 512)[::32]
    c0 = 3000 + 1500 * np.exp(-ff / 0.08)
    uf = j0(2 * np.pi * ff[None, :] * rr[:, None] / c0[None, :])
    uf = uf + rng.normal(0, 0.1, uf.shape) + 1j * rng.normal(0, 0.1, uf.shape)
    uf = uf.astype(np.complex64)
    U = uf.real.astype(np.float32)
    out["cfg1s"] = (U, rr.astype(np.float32), ff.astype(np.float32),
                    np.linspace(1000, 5000, 500)[::21].astype(np.float32))
    return out


def main():
    if not os.path.exists(LIB):
        sys.exit("build the reference first: make -C oracle ref")
    lib = ctypes.CDLL(LIB)
    arrays = {}
    for name, (U, r, f, c) in cases().items():
        arrays[name + "/U"] = U
        arrays[name + "/r"] = r
        arrays[name + "/f"] = f
        arrays[name + "/c"] = c
        for fn, tag in (("FJ", "fj"), ("FH", "fh")):
            for it in (0, 1):
                arrays[f"{name}/{tag}{it}"] = _call(lib, fn, U, r, f, c, it)
    path = os.path.join(HERE, "fjcpu_golden.npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
