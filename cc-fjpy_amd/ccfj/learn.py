"""Learned coarse-to-fine F-J: a small CNN proposes velocity bands.

No reference counterpart.  A coarse F-J spectrum (H = 64 velocity cells x
W = 256 frequencies) of a field is computed first; the CNN (HIP kernels,
bf16 weights, fp32 accumulation) marks where dispersion ridges pass; the
dense velocity grid (4000 values) is then evaluated only inside the proposed
bands (dilated by one coarse cell), and a frequency with no proposal falls
back to its full column.  The full-grid F-J stays the judge: ridge picks
(argmax over velocity per frequency) of the banded result are compared with
the full grid's.

The coarse spectrum of a cell is the maximum of |F-J| over the
sub-velocities inside it: the normalised transform on H S = 512 velocities
spaced evenly in slowness 1/c, max-pooled into the 64 cells (even in c).  A
ridge is about c^2 / (f R) wide in velocity (1/(f R) in slowness): 5 m/s at
1000 m/s and 1 Hz over 200 km, 125 m/s at 5000 m/s, against 63 m/s cells.
Even slowness steps sample it alike everywhere (3.6 samples a ridge width
at 1 Hz).  Even velocity steps (4 a cell) missed the slow, narrow ridges at
high frequency, whose columns then lost their picks; with 256 samples even
in slowness the ranking of near-equal features in noisy high-frequency
columns still failed now and then.  Measured on 64 fresh fields (ridge
agreement overall / in the worst field, evaluated fraction): 4 even-c
samples a cell 0.9984 / 0.930 at 13.9%; 256 even-slowness 0.9988 / 0.984 at
14.9%; 512 even-slowness 1.0 / 1.0 at 14.9%.  Cells between two samples (at
the fast end) take the two nearest.  The proposal is the network's mask
united with each column's coarse maximum and the cells within K_TOP of it.

Training data are generated on the device: synthetic fields with 1-4
modes (ccfj_synth_fields), their coarse F-J by one batched persistent
launch, and labels: cells within one cell of a dispersion curve whose
amplitude is at least 0.15 (ccfj_synth_labels), plus the cell of each
column's coarse maximum (ccfj_label_argmax).  Everything stays in device
memory (DeviceArray); the host only draws the random field parameters.

Choices (the workload description leaves them open): velocities
1000-5000 m/s, frequencies 0.01-1 Hz, 300 pairs with r ~ U[5, 200] km,
mode curves c_m(f) = cinf_m + d_m exp(-f/phi_m) with cinf_1 ~ U[1500, 3000],
each higher mode U[300, 900] m/s faster, d ~ U[500, 2500], phi ~ U[0.05,
0.5] Hz, amplitudes 1 then U[0.2, 0.8]; noise sigma = 0.06 sqrt(sum A^2)
10^(-SNR/20) with SNR ~ U[0, 20] dB.
"""
import time

import numpy as np

from . import _ccfj, _context, MODE_J

C_MIN, C_MAX = 1000.0, 5000.0
F_MIN, F_MAX = 0.01, 1.0
MAX_MODES = 4
K_TOP = 0.85   # cells whose coarse value reaches K_TOP x the column maximum are proposed


def pooling_ranges(c_sub, c_cells):
    """Sub-velocity index ranges [lo, hi) per cell (c_sub ascending): the
    samples inside the cell's half-open interval around its centre, or,
    where none falls inside, the two samples either side of its centre."""
    c_sub = np.asarray(c_sub, np.float64)
    c_cells = np.asarray(c_cells, np.float64)
    dc = c_cells[1] - c_cells[0]
    lo = np.searchsorted(c_sub, c_cells - dc / 2, side="left")
    hi = np.searchsorted(c_sub, c_cells + dc / 2, side="left")
    empty = hi <= lo
    mid = np.searchsorted(c_sub, c_cells, side="left")
    lo = np.where(empty, np.clip(mid - 1, 0, len(c_sub) - 2), lo)
    hi = np.where(empty, lo + 2, hi)
    return np.stack([lo, hi], axis=1).astype(np.int32)


def mode_windows(modes, f, c_dense, dc):
    """Dense-velocity windows [lo, hi) around every mode's curve: for each
    mode of amplitude > 0 (modes [MAX_MODES, 4] = (cinf, d, phi, amp),
    c_m(f) = cinf + d exp(-f / phi)) and each frequency index whose c_m(f)
    lies on [C_MIN, C_MAX], the indices of the ascending dense grid c_dense
    within dc of c_m(f).  Returns int arrays (frequency, lo, hi)."""
    f = np.asarray(f, np.float64)
    c_dense = np.asarray(c_dense)
    fi, lo, hi = [], [], []
    for cinf, d, phi, amp in np.asarray(modes, np.float64):
        if amp <= 0.0:
            continue
        cm = cinf + d * np.exp(-f / phi)
        i = np.nonzero((cm >= C_MIN) & (cm <= C_MAX))[0]
        fi.append(i)
        lo.append(np.searchsorted(c_dense, cm[i] - dc, side="left"))
        hi.append(np.searchsorted(c_dense, cm[i] + dc, side="right"))
    if not fi:
        return np.zeros(0, int), np.zeros(0, int), np.zeros(0, int)
    return np.concatenate(fi), np.concatenate(lo), np.concatenate(hi)


def sample_fields(rng, B, nr=300):
    """Host-side random parameters of B fields: r [B, nr] (sorted, metres),
    modes [B, 4, 4] (cinf, d, phi, amplitude; amplitude 0 = unused),
    sigma [B]."""
    r = np.sort(rng.uniform(5e3, 200e3, (B, nr)), axis=1).astype(np.float32)
    modes = np.zeros((B, MAX_MODES, 4), np.float32)
    nm = rng.integers(1, MAX_MODES + 1, B)
    for b in range(B):
        cinf = rng.uniform(1500.0, 3000.0)
        for m in range(nm[b]):
            modes[b, m] = (cinf, rng.uniform(500.0, 2500.0), rng.uniform(0.05, 0.5),
                           1.0 if m == 0 else rng.uniform(0.2, 0.8))
            cinf += rng.uniform(300.0, 900.0)
    snr = rng.uniform(0.0, 20.0, B)
    amp = np.sqrt(np.sum(modes[:, :, 3] ** 2, axis=1))
    sigma = (0.06 * amp * 10.0 ** (-snr / 20.0)).astype(np.float32)
    return r, modes, sigma


class Coarse2Fine:
    """Band-proposal CNN plus the coarse and dense F-J it steers."""

    def __init__(self, num=-1, nr=300, H=64, W=256, nc_dense=4000, max_batch=64, seed=0,
                 itype=1, sub=8):
        self.ctx = _context(num)
        self.nr, self.H, self.W, self.nc_dense = nr, H, W, nc_dense
        self.S = sub
        self.itype = itype
        self.max_batch = max_batch
        self.f = np.linspace(F_MIN, F_MAX, W).astype(np.float32)
        self.c_coarse = np.linspace(C_MIN, C_MAX, H).astype(np.float32)
        dc = (C_MAX - C_MIN) / (H - 1)
        # H S sub-velocities evenly spaced in slowness over the cells' span
        s_hi, s_lo = 1.0 / (C_MIN - dc / 2), 1.0 / (C_MAX + dc / 2)
        self.c_sub = (1.0 / np.linspace(s_hi, s_lo, H * sub)).astype(np.float32)
        self.ranges = pooling_ranges(self.c_sub, self.c_coarse)
        self.d_rg = _ccfj.device_array(self.ctx, self.ranges, np.int32)
        self.c_dense = np.linspace(C_MIN, C_MAX, nc_dense).astype(np.float32)
        self.d_f = _ccfj.device_array(self.ctx, self.f)
        self.d_cc = _ccfj.device_array(self.ctx, self.c_coarse)
        self.d_cs = _ccfj.device_array(self.ctx, self.c_sub)
        self.d_cd = _ccfj.device_array(self.ctx, self.c_dense)
        self.net = _ccfj.CNN(self.ctx, H, W, max_batch, seed)
        self._plans = {}

    # ------------------------------------------------------------ data
    def _fields(self, r, modes, sigma, seed):
        """Device fields [B*nr, W] for host parameters."""
        B = len(r)
        d_r = _ccfj.device_array(self.ctx, r)
        d_m = _ccfj.device_array(self.ctx, modes)
        d_s = _ccfj.device_array(self.ctx, sigma)
        U = _ccfj.DeviceArray(self.ctx, (B * self.nr, self.W))
        _ccfj.synth_fields(self.ctx, B, self.nr, self.W, d_r.addr, self.d_f.addr, d_m.addr,
                           d_s.addr, seed, U.addr)
        return U, d_r, d_m

    def _coarse(self, U, d_r, B, X, x0):
        """Coarse spectra of B fields into X rows x0..: the batched F-J on
        the H*S sub-velocities (normalised per column), max-pooled into the
        cells."""
        key = B
        if key not in self._plans:
            soff = np.arange(B + 1, dtype=np.int64) * self.nr
            self._plans[key] = (self.ctx.batch_plan(MODE_J, self.itype, True, soff, self.W,
                                                    self.H * self.S),
                                _ccfj.DeviceArray(self.ctx, (B, self.H * self.S, self.W)))
        plan, tmp = self._plans[key]
        plan.run(U.addr, U.addr, d_r.addr, self.d_f.addr, self.d_cs.addr, tmp.addr)
        _ccfj.pool_abs_max_ranges(self.ctx, tmp.addr, B, self.H * self.S, self.H, self.W,
                                  self.d_rg.addr, X.offset(x0 * self.H * self.W))

    def make_set(self, n, seed, chunk=1024):
        """n training examples on the device: X [n, H, W] coarse spectra,
        Y [n, H, W] ridge labels.  Returns (X, Y, host parameters)."""
        rng = np.random.default_rng(seed)
        X = _ccfj.DeviceArray(self.ctx, (n, self.H, self.W))
        Y = _ccfj.DeviceArray(self.ctx, (n, self.H, self.W))
        params = []
        for a in range(0, n, chunk):
            B = min(chunk, n - a)
            r, modes, sigma = sample_fields(rng, B, self.nr)
            U, d_r, d_m = self._fields(r, modes, sigma, seed * 1000003 + a)
            self._coarse(U, d_r, B, X, a)
            _ccfj.synth_labels(self.ctx, B, self.H, self.W, self.d_cc.addr, self.d_f.addr,
                               d_m.addr, float(self.c_coarse[1] - self.c_coarse[0]), 0.15,
                               Y.offset(a * self.H * self.W))
            _ccfj.label_argmax(self.ctx, X.offset(a * self.H * self.W), B, self.H, self.W,
                               Y.offset(a * self.H * self.W))
            params.append((r, modes, sigma))
        self.ctx.synchronize()
        return X, Y, params

    # ------------------------------------------------------------ training
    def train(self, X, Y, n, epochs=1, batch=None, lr=2e-3, seed=0, log=None):
        """Adam on shuffled mini-batches (the shuffle permutes batch starts
        of contiguous blocks).  Returns the per-step losses."""
        batch = batch or self.max_batch
        rng = np.random.default_rng(seed)
        losses = []
        starts = np.arange(0, n - batch + 1, batch)
        hw = self.H * self.W
        for ep in range(epochs):
            for k, s0 in enumerate(rng.permutation(starts)):
                losses.append(self.net.train_step(X.offset(int(s0) * hw), Y.offset(int(s0) * hw),
                                                  batch, lr))
            if log:
                log(f"epoch {ep + 1}: mean loss {np.mean(losses[-len(starts):]):.4f}")
        return np.array(losses)

    def logits(self, X, n):
        hw = self.H * self.W
        out = _ccfj.DeviceArray(self.ctx, (n, self.H, self.W))
        for a in range(0, n, self.max_batch):
            B = min(self.max_batch, n - a)
            self.net.forward(X.offset(a * hw), B, out.offset(a * hw))
        return out.download()

    # ------------------------------------------------------------ inference
    def items_from_mask(self, mask, dilate=1):
        """Items (frequency, first dense velocity, count <= 256) covering the
        coarse cells of mask [H, W] (dilated by `dilate` cells along velocity);
        a frequency with no cell falls back to its full column.  Returns
        (items int32 [n, 3], dense terms per station)."""
        dc = self.c_coarse[1] - self.c_coarse[0]
        items = []
        for i in range(self.W):
            on = mask[:, i]
            if dilate:
                on = np.convolve(on.astype(int), np.ones(2 * dilate + 1, int), "same") > 0
            if not on.any():
                lo, hi = np.array([0]), np.array([self.nc_dense])   # fallback: full column
            else:
                edges = np.diff(np.concatenate([[0], on.astype(int), [0]]))
                h0, h1 = np.nonzero(edges == 1)[0], np.nonzero(edges == -1)[0]
                cl = self.c_coarse[h0] - dc / 2
                ch = self.c_coarse[h1 - 1] + dc / 2
                lo = np.searchsorted(self.c_dense, cl)
                hi = np.searchsorted(self.c_dense, ch, side="right")
            for a, b in zip(lo, hi):
                for s in range(int(a), int(b), 256):
                    items.append((i, s, min(256, int(b) - s)))
        items = np.array(items, np.int32).reshape(-1, 3)
        return items, int(items[:, 2].sum())

    def argmax_mask(self, X):
        """Cell of each column's coarse maximum (X [H, W])."""
        m = np.zeros(X.shape, bool)
        m[np.argmax(np.abs(X), axis=0), np.arange(X.shape[1])] = True
        return m

    def top_mask(self, X, top=None):
        """Cells within `top` (default K_TOP) of their column's coarse
        maximum (X [H, W], the maximum's own cell included)."""
        a = np.abs(X)
        return a >= (K_TOP if top is None else top) * np.max(a, axis=0, keepdims=True)

    def propose_items(self, logit, X, thresh=0.0, dilate=1, top=None):
        """Bands where the network's logit exceeds thresh (0: p > 0.5),
        united with the cells within `top` of each column's coarse maximum
        (X [H, W]): near-ties between ridges, and noise-dominated columns,
        are then evaluated wherever their maximum may lie."""
        return self.items_from_mask((logit > thresh) | self.argmax_mask(X) | self.top_mask(X, top),
                                    dilate)

    def coarse_items(self, X, level=0.5, dilate=1):
        """The same banding without the network: coarse cells whose
        normalised |F-J| reaches `level`."""
        return self.items_from_mask(np.abs(X) >= level, dilate)

    def dense(self, U_rows, d_r_rows, items=None):
        """Dense F-J [nc_dense, W] of one field (device rows of U and r):
        the full grid, or only the listed items (zeros elsewhere)."""
        out = _ccfj.DeviceArray(self.ctx, (self.nc_dense, self.W))
        if items is None:
            self.ctx.transform_dev(MODE_J, self.itype, False, U_rows, U_rows, d_r_rows,
                                   self.d_f.addr, self.d_cd.addr, self.nr, self.W, self.nc_dense,
                                   out.addr, 0)
        else:
            out.upload(np.zeros((self.nc_dense, self.W), np.float32))
            d_it = _ccfj.device_array(self.ctx, items, np.int32)
            _ccfj.transform_items(self.ctx, MODE_J, self.itype, U_rows, U_rows, d_r_rows,
                                  self.d_f.addr, self.d_cd.addr, self.nr, self.W, self.nc_dense,
                                  d_it.addr, len(items), out.addr)
        self.ctx.synchronize()
        return out.download()

    def mode_windows(self, modes):
        """mode_windows on this model's grids."""
        return mode_windows(modes, self.f, self.c_dense, self.c_coarse[1] - self.c_coarse[0])

    @staticmethod
    def mode_picks(spec, windows):
        """|spec| argmax over velocity inside each mode window (spec
        [nc_dense, W]) -> dense velocity indices."""
        fi, lo, hi = windows
        a = np.abs(spec)
        return np.array([lo[k] + int(np.argmax(a[lo[k]:hi[k], fi[k]])) for k in range(len(fi))],
                        np.int64)

    def evaluate(self, n, seed, thresh=0.0, dilate=1, levels=(), use_net=True):
        """Fresh fields: coarse F-J -> bands -> banded dense F-J, against the
        full grid.  Bands come from the network (method "cnn", unless
        use_net is False) and from thresholding the coarse spectrum at each
        of `levels` (methods "coarse@level").  Per method: ridge agreement
        (fraction of frequencies whose |F-J| argmax over velocity is within
        one dense step of the full grid's), fraction of dense terms
        evaluated, and main-kernel time (HIP events); plus the full grid's
        kernel time and the wall time of the coarse F-J and of the network
        for all n fields.  Mode agreement: the same one-step test for every
        mode of the field, on the |F-J| argmax within one coarse cell of the
        mode's curve (mode_windows), so that weaker modes, whose ridges the
        global argmax never picks, count too."""
        rng = np.random.default_rng(seed)
        r, modes, sigma = sample_fields(rng, n, self.nr)
        U, d_r, _ = self._fields(r, modes, sigma, seed * 7919 + 1)
        X = _ccfj.DeviceArray(self.ctx, (n, self.H, self.W))
        self._coarse(U, d_r, n, X, 0)     # first call: plan, tuning, workspace
        self.ctx.synchronize()
        t0 = time.perf_counter()
        self._coarse(U, d_r, n, X, 0)
        self.ctx.synchronize()
        t1 = time.perf_counter()
        Xh = X.download()
        t2 = time.perf_counter()
        lg = self.logits(X, n) if use_net else None
        t3 = time.perf_counter()
        methods = (["cnn"] if use_net else []) + [f"coarse@{v:g}" for v in levels]
        stats = {m: {"agree": [], "frac": [], "mode": [], "ms": 0.0} for m in methods}
        full_ms = 0.0
        self.ctx.set_timing(True)
        self.ctx.get_timing()
        try:
            for b in range(n):
                rows = U.offset(b * self.nr * self.W)
                rb = d_r.offset(b * self.nr)
                full = self.dense(rows, rb)
                full_ms += self.ctx.get_timing()[0]
                pf = np.argmax(np.abs(full), axis=0)
                win = self.mode_windows(modes[b])
                mf = self.mode_picks(full, win)
                for m in methods:
                    if m == "cnn":
                        items, terms = self.propose_items(lg[b], Xh[b], thresh, dilate)
                    else:
                        items, terms = self.coarse_items(Xh[b], float(m.split("@")[1]), dilate)
                    band = self.dense(rows, rb, items)
                    stats[m]["ms"] += self.ctx.get_timing()[0]
                    pb = np.argmax(np.abs(band), axis=0)
                    stats[m]["agree"].append(np.mean(np.abs(pf - pb) <= 1))
                    if len(mf):
                        stats[m]["mode"].append(np.mean(np.abs(mf - self.mode_picks(band, win)) <= 1))
                    stats[m]["frac"].append(terms / (self.nc_dense * self.W))
        finally:
            self.ctx.set_timing(False)
        out = {"n": n, "full_kernel_ms": full_ms, "coarse_ms": (t1 - t0) * 1e3,
               "cnn_ms": (t3 - t2) * 1e3,
               "coarse_fraction": self.H * self.S / self.nc_dense}
        for m, v in stats.items():
            out[m] = {"ridge_agreement": float(np.mean(v["agree"])),
                      "worst_field_agreement": float(np.min(v["agree"])),
                      "mode_agreement": float(np.mean(v["mode"])) if v["mode"] else 1.0,
                      "worst_field_mode_agreement": float(np.min(v["mode"])) if v["mode"] else 1.0,
                      "evaluated_fraction": float(np.mean(v["frac"])),
                      "kernel_ms": v["ms"]}
        return out

    def save(self, path):
        np.save(path, self.net.get_params())

    def load(self, path):
        self.net.set_params(np.load(path, allow_pickle=False))
