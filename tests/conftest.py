import os
import sys

import numpy as np
import pytest

# torch, when present, must bring up its HIP runtime before libccfj.so is
# loaded, so that both share one runtime in this process
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "cc-fjpy_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "fjcpu_golden.npz")
GOLDEN_CASES = ["nr2", "nr3", "edges17", "branches", "cfg1s"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libccfj.so")


@pytest.fixture(scope="session")
def golden():
    d = np.load(GOLDEN)
    return {k: d[k] for k in d.files}


def golden_case(g, name):
    return tuple(g[f"{name}/{k}"] for k in "Urfc")


@pytest.fixture(scope="session")
def ccfj():
    """The package with its HIP library loaded and a GPU visible."""
    import ccfj as m
    assert m.device_count() >= 1, "no GPU visible"
    return m


@pytest.fixture(scope="session")
def ctx(ccfj):
    return ccfj._context(-1)
