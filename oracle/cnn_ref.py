"""CPU restatement of the band-proposal CNN.  TEST INFRASTRUCTURE ONLY.

float64 numpy forward pass and hand-derived backward pass of
conv3x3(1->32)+ReLU, conv3x3(32->32)+ReLU, conv3x3(32->32)+ReLU,
conv3x3(32->1), mean binary cross-entropy with logits -- the network of
cc-fjpy_amd/csrc/fj_cnn.hip, with the same flat parameter layout
(w1 b1 w2 b2 w3 b3 w4 b4).  Weights are rounded to bfloat16 (round to
nearest even) before use, as the GPU kernels read them; biases are not
(round_weights=False keeps them exact, which the finite-difference check of
this restatement needs: with rounding the loss is piecewise constant in the
weights).  The gradient is taken w.r.t. the weights as used, and applied to
the fp32 master copy unchanged.
"""
import numpy as np

C = 32
LAYERS = [(1, C), (C, C), (C, C), (C, 1)]


def split(params):
    out, o = [], 0
    for ci, co in LAYERS:
        w = params[o:o + co * ci * 9].reshape(co, ci, 3, 3)
        o += co * ci * 9
        b = params[o:o + co]
        o += co
        out.append((w, b))
    return out


def bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def conv(x, w, b):
    B, CI, H, W = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    y = np.zeros((B, w.shape[0], H, W)) + b[None, :, None, None]
    for kh in range(3):
        for kw in range(3):
            y += np.einsum("bchw,oc->bohw", xp[:, :, kh:kh + H, kw:kw + W], w[:, :, kh, kw])
    return y


def conv_back(x, w, dz):
    """Gradients of y = conv(x, w, b) given dz = dL/dy: (dx, dw, db)."""
    B, CI, H, W = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    dzp = np.pad(dz, ((0, 0), (0, 0), (1, 1), (1, 1)))
    dw = np.zeros_like(w)
    dx = np.zeros_like(x)
    for kh in range(3):
        for kw in range(3):
            dw[:, :, kh, kw] = np.einsum("bohw,bchw->oc", dz, xp[:, :, kh:kh + H, kw:kw + W])
            dx += np.einsum("bohw,oc->bchw", dzp[:, :, 2 - kh:2 - kh + H, 2 - kw:2 - kw + W],
                            w[:, :, kh, kw])
    return dx, dw, dz.sum(axis=(0, 2, 3))


def _w(w, round_weights):
    return bf16(w) if round_weights else w


def forward(params, x, round_weights=True):
    """x [B, H, W] -> logits [B, H, W] and the activations."""
    a = np.asarray(x, np.float64)[:, None]
    acts = [a]
    for l, (w, b) in enumerate(split(np.asarray(params, np.float64))):
        z = conv(a, _w(w, round_weights), b)
        a = np.maximum(z, 0.0) if l < len(LAYERS) - 1 else z
        acts.append(a)
    return a[:, 0], acts


def loss(params, x, t, round_weights=True):
    z, _ = forward(params, x, round_weights)
    t = np.asarray(t, np.float64)
    return np.sum(np.maximum(z, 0) - z * t + np.log1p(np.exp(-np.abs(z)))) / z.size


def loss_and_grads(params, x, t, round_weights=True):
    """Mean BCE-with-logits loss and its gradient (flat, parameter layout)."""
    params = np.asarray(params, np.float64)
    z, acts = forward(params, x, round_weights)
    t = np.asarray(t, np.float64)
    n = z.size
    value = np.sum(np.maximum(z, 0) - z * t + np.log1p(np.exp(-np.abs(z)))) / n
    g = ((1.0 / (1.0 + np.exp(-z)) - t) / n)[:, None]
    grads = []
    layers = split(params)
    for l in range(len(LAYERS) - 1, -1, -1):
        w, b = layers[l]
        if l < len(LAYERS) - 1:
            g = g * (acts[l + 1] > 0)
        dx, dw, db = conv_back(acts[l], _w(w, round_weights), g)
        grads.append((dw, db))
        g = dx
    flat = []
    for dw, db in reversed(grads):
        flat += [dw.ravel(), db.ravel()]
    return value, np.concatenate(flat)
