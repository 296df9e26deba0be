"""The C ABI library without a GPU: it loads, exports every entry point
include/ccfj.h declares, and reports errors by return code (never by
exiting the process, unlike the reference's GPU helper,
include/cuda_helper.h:74-84)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ccfj.h")
LIB = os.path.join(ROOT, "cc-fjpy_amd", "ccfj", "libccfj.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        import __graft_entry__
        __graft_entry__.build()
    return ctypes.CDLL(LIB)


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    for required in ("FJ", "FHr", "FHi", "FH", "ccfj_create", "ccfj_transform",
                     "ccfj_transform_dev", "ccfj_plain_fp64", "ccfj_last_error"):
        assert required in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _no_gpu(lib):
    lib.ccfj_device_count.restype = ctypes.c_int
    return lib.ccfj_device_count() == 0


def test_errors_without_gpu(lib):
    if not _no_gpu(lib):
        pytest.skip("a GPU is visible")
    lib.ccfj_last_error.restype = ctypes.c_char_p
    ctx = ctypes.c_void_p()
    assert lib.ccfj_create(-1, ctypes.byref(ctx)) == -2          # CCFJ_EDEVICE
    assert b"no HIP GPU" in lib.ccfj_last_error()
    assert not ctx.value
    U = np.ones(12, np.float32)
    out = np.zeros(20, np.float32)
    r = np.arange(3, dtype=np.float32)
    f = np.arange(4, dtype=np.float32)
    c = np.arange(1, 6, dtype=np.float32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib.FJ.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int] * 4
    rc = lib.FJ(p(U), p(r), p(f), p(out), p(c), 5, 3, 4, 1)
    assert rc != 0 and np.all(out == 0)


def test_null_context_and_bad_arguments(lib):
    lib.ccfj_last_error.restype = ctypes.c_char_p
    lib.ccfj_transform.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 3 + [ctypes.c_void_p] * 5 + [ctypes.c_int] * 3 + [ctypes.c_void_p]
    assert lib.ccfj_transform(None, 0, 0, 0, None, None, None, None, None, 1, 1, 1, None) == -1
    assert b"null context" in lib.ccfj_last_error()
    assert lib.ccfj_destroy(None) == 0


def test_cython_shim_imports_and_fails_loudly_without_gpu():
    import ccfj
    from ccfj import _ccfj
    assert callable(_ccfj.fj_full) and callable(_ccfj.fHi_full)
    if ccfj.device_count() == 0:
        with pytest.raises(RuntimeError, match="no HIP GPU"):
            _ccfj.Context(-1)
        with pytest.raises(RuntimeError):
            ccfj.fj_noise(np.ones((4, 6)), np.arange(4.) * 1e3, np.array([1e3, 2e3]),
                          np.linspace(0.1, 0.6, 6))
