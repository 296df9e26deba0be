"""ccfj_transform_dev_noalloc through the C ABI: the same results as
ccfj_transform_dev, bit for bit, on two calls in a row (the second on a
smaller problem, so stale work-array contents would show), for both
integration types and single- and two-spectrum modes; a problem larger than
its fixed work array is refused with the bytes it would need."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

P, I = ctypes.c_void_p, ctypes.c_int
MODE_J, MODE_Y, MODE_J_MINUS_Y = 0, 1, 2


@pytest.fixture(scope="module")
def lib():
    import torch  # noqa: F401  (torch's HIP runtime first, as the package does)
    L = ctypes.CDLL(os.path.join(ROOT, "cc-fjpy_amd", "ccfj", "libccfj.so"))
    sig = [P, I, I, I] + [P] * 5 + [I] * 3 + [P, P]
    for name in ("ccfj_transform_dev", "ccfj_transform_dev_noalloc"):
        getattr(L, name).argtypes = sig
        getattr(L, name).restype = I
    L.ccfj_create.argtypes = [I, ctypes.POINTER(P)]
    L.ccfj_create.restype = I
    L.ccfj_destroy.argtypes = [P]
    L.ccfj_last_error.restype = ctypes.c_char_p
    return L


@pytest.fixture(scope="module")
def ctx(lib):
    h = P()
    assert lib.ccfj_create(0, ctypes.byref(h)) == 0, lib.ccfj_last_error()
    yield h
    lib.ccfj_destroy(h)


def _problem(seed, nr, nf, nc):
    import torch
    rng = np.random.default_rng(seed)
    r = np.sort(rng.uniform(2e3, 150e3, nr)).astype(np.float32)
    f = np.linspace(0.01, 0.8, nf).astype(np.float32)
    c = np.linspace(900.0, 5200.0, nc).astype(np.float32)
    U0 = rng.standard_normal((nr, nf)).astype(np.float32)
    U1 = rng.standard_normal((nr, nf)).astype(np.float32)
    dev = torch.device("cuda:0")
    return {k: torch.from_numpy(v).to(dev) for k, v in
            dict(U0=U0, U1=U1, r=r, f=f, c=c).items()}, (nr, nf, nc)


def _call(lib, name, ctx, mode, itype, norm, t, dims, out, stream):
    nr, nf, nc = dims
    rc = getattr(lib, name)(ctx, mode, itype, norm, t["U0"].data_ptr(), t["U1"].data_ptr(),
                            t["r"].data_ptr(), t["f"].data_ptr(), t["c"].data_ptr(), nr, nf, nc,
                            out.data_ptr(), stream)
    return rc


@pytest.mark.parametrize("mode,itype,norm", [(MODE_J, 1, 1), (MODE_J_MINUS_Y, 1, 1),
                                             (MODE_J, 0, 0), (MODE_Y, 0, 1),
                                             (MODE_J_MINUS_Y, 0, 0)])
def test_noalloc_matches_transform_dev(lib, ctx, mode, itype, norm):
    import torch
    s = torch.cuda.Stream(device=torch.device("cuda:0"))
    # 300 pairs x 512 f x 500 c, then a smaller problem on the same array
    for seed, (nr, nf, nc) in ((1, (300, 512, 500)), (2, (57, 40, 130)), (3, (1, 9, 20))):
        t, dims = _problem(seed, nr, nf, nc)
        want = torch.full((nc, nf), float("nan"), device="cuda:0")
        got = torch.full((nc, nf), float("nan"), device="cuda:0")
        torch.cuda.synchronize()
        print(f"noalloc mode={mode} itype={itype} norm={norm} nr={nr} nf={nf} nc={nc}", flush=True)
        assert _call(lib, "ccfj_transform_dev", ctx, mode, itype, norm, t, dims, want,
                     s.cuda_stream) == 0, lib.ccfj_last_error()
        assert _call(lib, "ccfj_transform_dev_noalloc", ctx, mode, itype, norm, t, dims, got,
                     s.cuda_stream) == 0, lib.ccfj_last_error()
        s.synchronize()
        a, b = want.cpu().numpy(), got.cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (nr, nf, nc)
        if nr < 2:
            assert np.all(a == 0) or norm       # empty sums (NaN once normalised)
        else:
            assert np.all(np.isfinite(a))


def test_noalloc_refuses_oversize(lib, ctx):
    import torch
    dev = torch.device("cuda:0")
    nr, nf, nc = 19900, 512, 3000          # about 330 MB of records
    x = torch.zeros(16, device=dev)
    rc = lib.ccfj_transform_dev_noalloc(ctx, MODE_J_MINUS_Y, 1, 1, x.data_ptr(), x.data_ptr(),
                                        x.data_ptr(), x.data_ptr(), x.data_ptr(), nr, nf, nc,
                                        x.data_ptr(), None)
    assert rc == -1
    msg = lib.ccfj_last_error().decode()
    assert "needs" in msg and "67108864" in msg, msg
