"""The CPU restatement of the band-proposal CNN (oracle/cnn_ref.py), pinned
independently of itself: the forward convolution against scipy's 2-D
correlation, the backward pass against central finite differences over
every parameter block, and the bf16 rounding against torch's conversion."""
import numpy as np
import pytest
import scipy.signal

from oracle import cnn_ref as C


def _params(rng, bias_scale=0.1):
    blocks = []
    for ci, co in C.LAYERS:
        blocks.append(rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), co * ci * 9))
        blocks.append(rng.normal(0.0, bias_scale, co))
    return np.concatenate(blocks)


def _block_ranges():
    out, o = [], 0
    for ci, co in C.LAYERS:
        out.append(("w", o, o + co * ci * 9))
        o += co * ci * 9
        out.append(("b", o, o + co))
        o += co
    return out


def test_parameter_count_matches_layout():
    assert sum(co * ci * 9 + co for ci, co in C.LAYERS) == 19_105
    assert _block_ranges()[-1][2] == 19_105


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 3, 7, 9))
    w = rng.normal(size=(4, 3, 3, 3))
    b = rng.normal(size=4)
    got = C.conv(x, w, b)
    want = np.zeros_like(got)
    for n in range(2):
        for o in range(4):
            want[n, o] = b[o] + sum(scipy.signal.correlate2d(x[n, i], w[o, i], mode="same")
                                    for i in range(3))
    assert np.max(np.abs(got - want)) < 1e-12


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    p = _params(rng)
    x = rng.normal(size=(2, 5, 6))
    t = (rng.random((2, 5, 6)) > 0.7).astype(np.float64)
    _, g = C.loss_and_grads(p, x, t, round_weights=False)
    eps = 1e-6
    for name, a, b in _block_ranges():
        idx = rng.choice(np.arange(a, b), size=min(6, b - a), replace=False)
        for i in idx:
            e = np.zeros_like(p)
            e[i] = eps
            fd = (C.loss(p + e, x, t, False) - C.loss(p - e, x, t, False)) / (2 * eps)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, np.max(np.abs(g))), (name, i, fd, g[i])


def test_loss_is_mean_bce():
    rng = np.random.default_rng(3)
    p = _params(rng)
    x = rng.normal(size=(1, 4, 4))
    t = (rng.random((1, 4, 4)) > 0.5).astype(np.float64)
    z, _ = C.forward(p, x)
    s = 1.0 / (1.0 + np.exp(-z))
    want = -np.mean(t * np.log(s) + (1 - t) * np.log(1 - s))
    assert abs(C.loss(p, x, t) - want) < 1e-12


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    a = np.concatenate([rng.normal(size=1000).astype(np.float32),
                        np.array([1.0, -2.5, 3.0e-3, 65504.0, 1.00390625, 1.01171875],
                                 np.float32)])
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(C.bf16(a).astype(np.float32), want)
