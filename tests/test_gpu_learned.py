"""Learned coarse-to-fine path on the GPU: the CNN's forward pass, gradients
and optimizer step against the CPU restatement (oracle/cnn_ref.py), the
device generators of training data against numpy restatements, and the
item-list F-J against the full grid."""
import numpy as np
import pytest
import scipy.special

from oracle import cnn_ref as CR

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(ccfj):
    from ccfj import _ccfj
    return _ccfj


def _params(rng):
    blocks = []
    for ci, co in CR.LAYERS:
        blocks.append(rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), co * ci * 9))
        blocks.append(rng.normal(0.0, 0.05, co))
    return np.concatenate(blocks).astype(np.float32)


def _blocks():
    o = 0
    for l, (ci, co) in enumerate(CR.LAYERS):
        yield f"w{l + 1}", o, o + co * ci * 9
        o += co * ci * 9
        yield f"b{l + 1}", o, o + co
        o += co


@pytest.mark.parametrize("B,H,W", [(3, 20, 40), (2, 64, 256)])
def test_cnn_forward_matches_oracle(lib, ctx, B, H, W):
    rng = np.random.default_rng(10 + H)
    net = lib.CNN(ctx, H, W, 4, 0)
    p = _params(rng)
    net.set_params(p)
    x = rng.normal(size=(B, H, W)).astype(np.float32)
    dx = lib.device_array(ctx, x)
    dz = lib.DeviceArray(ctx, (B, H, W))
    net.forward(dx.addr, B, dz.addr)
    got = dz.download()
    want, _ = CR.forward(p.astype(np.float64), x)
    err = np.max(np.abs(got - want)) / max(1.0, np.max(np.abs(want)))
    print(f"forward B={B} H={H} W={W}: max rel err {err:.2e}")
    assert err <= 2e-5


@pytest.mark.parametrize("B,H,W", [(3, 20, 40), (2, 64, 256)])
def test_cnn_gradients_match_oracle(lib, ctx, B, H, W):
    rng = np.random.default_rng(20 + H)
    net = lib.CNN(ctx, H, W, 4, 0)
    p = _params(rng)
    net.set_params(p)
    x = rng.normal(size=(B, H, W)).astype(np.float32)
    t = (rng.random((B, H, W)) > 0.8).astype(np.float32)
    dx, dt = lib.device_array(ctx, x), lib.device_array(ctx, t)
    g, loss = net.gradients(dx.addr, dt.addr, B)
    want_loss, want = CR.loss_and_grads(p.astype(np.float64), x, t)
    print(f"loss gpu {loss:.7f} oracle {want_loss:.7f}")
    assert abs(loss - want_loss) <= 1e-5 * max(1.0, want_loss)
    for name, a, b in _blocks():
        scale = np.max(np.abs(want[a:b]))
        err = np.max(np.abs(g[a:b] - want[a:b])) / scale
        print(f"  {name}: max |g| {scale:.3e}, max rel err {err:.2e}")
        assert err <= 1e-3, name


def test_cnn_adam_step(lib, ctx):
    """One training step = gradient (checked above) + Adam at step 1, where
    the bias-corrected update is lr * g / (|g| + eps)."""
    rng = np.random.default_rng(30)
    B, H, W = 2, 16, 32
    net = lib.CNN(ctx, H, W, 2, 0)
    p = _params(rng)
    net.set_params(p)
    x = rng.normal(size=(B, H, W)).astype(np.float32)
    t = (rng.random((B, H, W)) > 0.8).astype(np.float32)
    dx, dt = lib.device_array(ctx, x), lib.device_array(ctx, t)
    g, _ = net.gradients(dx.addr, dt.addr, B)
    lr = 1e-3
    net.train_step(dx.addr, dt.addr, B, lr)
    got = net.get_params()
    want = p - lr * g / (np.abs(g) + 1e-8)
    assert np.max(np.abs(got - want)) <= 1e-6


def _splitmix64(x):
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def _gauss(key):
    a = _splitmix64(key)
    b = _splitmix64(key ^ np.uint64(0xD1B54A32D192ED03))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 0.5) / 9007199254740992.0
    u2 = (b >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2)).astype(np.float32)


def test_synthetic_fields_and_labels(lib, ctx):
    from ccfj import learn
    rng = np.random.default_rng(40)
    B, nr, nf, H = 3, 50, 40, 16
    r, modes, sigma = learn.sample_fields(rng, B, nr)
    f = np.linspace(0.01, 1.0, nf).astype(np.float32)
    cg = np.linspace(1000, 5000, H).astype(np.float32)
    seed = 12345
    d = [lib.device_array(ctx, a) for a in (r, f, modes, sigma, cg)]
    U = lib.DeviceArray(ctx, (B * nr, nf))
    lib.synth_fields(ctx, B, nr, nf, d[0].addr, d[1].addr, d[2].addr, d[3].addr, seed, U.addr)
    Y = lib.DeviceArray(ctx, (B, H, nf))
    hw, amin = float(cg[1] - cg[0]), 0.15
    lib.synth_labels(ctx, B, H, nf, d[4].addr, d[1].addr, d[2].addr, hw, amin, Y.addr)
    got_U = U.download().reshape(B, nr, nf)
    got_Y = Y.download()
    b_, s_, i_ = np.meshgrid(np.arange(B), np.arange(nr), np.arange(nf), indexing="ij")
    key = np.uint64(seed) ^ (((b_.astype(np.uint64) * np.uint64(1000003) + s_.astype(np.uint64))
                              * np.uint64(1000033)) + i_.astype(np.uint64))
    noise = _gauss(key)
    fd = f.astype(np.float64)
    for b in range(B):
        sig = np.zeros((nr, nf))
        for m in range(learn.MAX_MODES):
            cinf, dd, phi, amp = (float(v) for v in modes[b, m])
            if amp == 0.0:
                continue
            cm = cinf + dd * np.exp(-fd / phi)
            sig += amp * scipy.special.j0(2 * np.pi * fd[None, :] * r[b, :, None].astype(np.float64)
                                          / cm[None, :])
        want = sig.astype(np.float32) + sigma[b] * noise[b]
        assert np.max(np.abs(got_U[b] - want)) <= 2e-6 * learn.MAX_MODES
        lab = np.zeros((H, nf), np.float32)
        close = np.zeros((H, nf), bool)
        for m in range(learn.MAX_MODES):
            cinf, dd, phi, amp = (float(v) for v in modes[b, m])
            if amp < amin:
                continue
            cm = cinf + dd * np.exp(-fd / phi)
            dist = np.abs(cg[:, None] - cm[None, :])
            lab[dist <= hw] = 1.0
            close |= np.abs(dist - hw) < 1e-2
        assert np.all((got_Y[b] == lab) | close)
        assert got_Y[b].sum() > 0


def test_items_match_full_grid(ccfj, lib, ctx):
    rng = np.random.default_rng(50)
    nr, nf, nc = 120, 48, 700
    r = np.sort(rng.uniform(5e3, 200e3, nr)).astype(np.float32)
    U = rng.normal(size=(nr, nf)).astype(np.float32)
    f = np.linspace(0.01, 1.0, nf).astype(np.float32)
    c = np.linspace(1000, 5000, nc).astype(np.float32)
    dU, dr, df, dc = (lib.device_array(ctx, a) for a in (U, r, f, c))
    full = lib.DeviceArray(ctx, (nc, nf))
    for itype in (0, 1):
        ctx.transform_dev(ccfj.MODE_J, itype, False, dU.addr, dU.addr, dr.addr, df.addr, dc.addr,
                          nr, nf, nc, full.addr, 0)
        ctx.synchronize()
        want = full.download()
        items = [(0, 0, 256), (5, 100, 37), (47, 444, 256), (20, 690, 256), (48, 0, 10)]
        items = np.array(items, np.int32)   # the last two are clipped / skipped in the kernel
        band = lib.device_array(ctx, np.zeros((nc, nf), np.float32))
        it = lib.device_array(ctx, items, np.int32)
        lib.transform_items(ctx, ccfj.MODE_J, itype, dU.addr, dU.addr, dr.addr, df.addr, dc.addr,
                            nr, nf, nc, it.addr, len(items), band.addr)
        ctx.synchronize()
        got = band.download()
        mask = np.zeros((nc, nf), bool)
        for fi, c0, n in items:
            if fi < nf:
                mask[c0:min(c0 + n, nc), fi] = True
        assert np.array_equal(got[mask], want[mask]), itype
        assert np.all(got[~mask] == 0)


def test_pool_and_argmax_labels(lib, ctx):
    rng = np.random.default_rng(60)
    B, H, S, W = 3, 8, 4, 40
    x = rng.normal(size=(B, H * S, W)).astype(np.float32)
    dx = lib.device_array(ctx, x)
    dp = lib.DeviceArray(ctx, (B, H, W))
    lib.pool_abs_max(ctx, dx.addr, B, H, S, W, dp.addr)
    pooled = dp.download()
    assert np.array_equal(pooled, np.abs(x).reshape(B, H, S, W).max(axis=2))
    y0 = (rng.random((B, H, W)) > 0.9).astype(np.float32)
    dy = lib.device_array(ctx, y0)
    lib.label_argmax(ctx, dp.addr, B, H, W, dy.addr)
    want = y0.copy()
    for b in range(B):
        want[b, np.argmax(pooled[b], axis=0), np.arange(W)] = 1.0
    assert np.array_equal(dy.download(), want)


def test_pool_ranges(lib, ctx):
    """Pooling over per-cell index ranges (the coarse cells over slowness-spaced
    sub-velocities), clipped to the input and 0 for an empty range."""
    from ccfj import learn
    rng = np.random.default_rng(61)
    B, N, H, W = 2, 37, 6, 33
    x = rng.normal(size=(B, N, W)).astype(np.float32)
    ranges = np.array([[0, 5], [5, 6], [6, 20], [19, 37], [30, 45], [10, 10]], np.int32)
    dx = lib.device_array(ctx, x)
    dr = lib.device_array(ctx, ranges, np.int32)
    dp = lib.DeviceArray(ctx, (B, H, W))
    lib.pool_abs_max_ranges(ctx, dx.addr, B, N, H, W, dr.addr, dp.addr)
    got = dp.download()
    for h, (a, b) in enumerate(ranges):
        b = min(b, N)
        want = np.abs(x[:, a:b]).max(axis=1) if b > a else np.zeros((B, W), np.float32)
        assert np.array_equal(got[:, h], want), h
    # the ranges the coarse-to-fine path uses cover every cell, in order
    m = learn.pooling_ranges(np.linspace(900.0, 5100.0, 300) ** 1.0, np.linspace(1000, 5000, 64))
    assert np.all(m[:, 1] > m[:, 0]) and np.all(np.diff(m[:, 0]) >= 0)


def test_trained_bands_beat_the_method_without_the_network():
    """What the trained network gains: on fresh fields its bands keep every
    full-grid ridge pick (the global |F-J| maximum of each frequency), and
    they also hold the weaker modes.  Thresholding the coarse spectrum (the
    same banding without the network) needs a much lower threshold, and far
    more dense terms, to keep as many picks on every mode's curve; the
    untrained network's bands do much worse."""
    from ccfj import learn
    m = learn.Coarse2Fine(num=-1, seed=0)
    untrained = m.evaluate(16, seed=11)
    X, Y, _ = m.make_set(50000, seed=1)        # the workload's 50,000 training spectra
    losses = m.train(X, Y, 50000, epochs=3, lr=2e-3)
    trained = m.evaluate(16, seed=11, levels=(0.5, 0.3))
    cnn, hi, lo, raw = (trained["cnn"], trained["coarse@0.5"], trained["coarse@0.3"],
                        untrained["cnn"])
    print(f"loss {losses[0]:.3f} -> {losses[-20:].mean():.3f}")
    for name, r in (("trained", cnn), ("coarse@0.5", hi), ("coarse@0.3", lo), ("untrained", raw)):
        print(f"{name:11s} agreement {r['ridge_agreement']:.4f} (worst field "
              f"{r['worst_field_agreement']:.3f}), modes {r['mode_agreement']:.4f} at "
              f"{r['evaluated_fraction']:.3f} of the terms")
    assert losses[-20:].mean() < 0.3 * losses[0]
    # ridge picks within one dense velocity step of the full grid's
    assert cnn["ridge_agreement"] >= 0.998
    assert cnn["worst_field_agreement"] >= 0.99
    # the modes: as many picks as the threshold at 0.3 at well under its
    # cost, and far more than the threshold at 0.5 at about its cost
    assert cnn["mode_agreement"] >= lo["mode_agreement"] - 0.02
    assert cnn["evaluated_fraction"] <= 0.8 * lo["evaluated_fraction"]
    assert cnn["mode_agreement"] >= hi["mode_agreement"] + 0.04
    assert cnn["evaluated_fraction"] <= 1.2 * hi["evaluated_fraction"]
    assert cnn["evaluated_fraction"] <= 0.5 * raw["evaluated_fraction"] or \
        cnn["mode_agreement"] >= raw["mode_agreement"] + 0.05
