"""Batched sub-array F-J on the GPU: one persistent launch over many uneven
sub-arrays must give what one fj_noise call per sub-array gives, meet the
FP64-oracle bar, and replay correctly from a captured HIP graph."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def batch():
    from ccfj import synth
    d = synth.subarrays(seed=11, nb=37, nmin=2, nmax=90, nf=40, nc=150)
    # edge members: a single station and a two-station sub-array
    d["ufs"][3] = d["ufs"][3][:1]
    d["rs"][3] = d["rs"][3][:1]
    d["ufs"][5] = d["ufs"][5][:2]
    d["rs"][5] = d["rs"][5][:2]
    return d


@pytest.mark.parametrize("itype", [0, 1])
@pytest.mark.parametrize("func", [0, 1])
def test_batch_equals_per_subarray_calls(ccfj, batch, itype, func):
    d = batch
    got = ccfj.fj_noise_batch(d["ufs"], d["rs"], d["c"], d["f"], itype=itype, func=func)
    assert got.shape == (len(d["ufs"]), len(d["c"]), len(d["f"])) and got.dtype == np.float32
    worst = 0.0
    for b, (uf, r) in enumerate(zip(d["ufs"], d["rs"])):
        with np.errstate(invalid="ignore", divide="ignore"):
            one = ccfj.fj_noise(uf, r, d["c"], d["f"], itype=itype, func=func)
        assert np.array_equal(np.isnan(one), np.isnan(got[b]))
        fin = np.isfinite(one)
        worst = max(worst, float(np.max(np.abs(one[fin] - got[b][fin]), initial=0.0)))
    assert worst <= 1e-6


@pytest.mark.parametrize("itype", [0, 1])
def test_batch_vs_oracle(ccfj, batch, itype):
    d = batch
    got = ccfj.fj_noise_batch(d["ufs"], d["rs"], d["c"], d["f"], itype=itype, func=0)
    for b in (0, 7, 20, 36):
        want = O.fj_noise(d["ufs"][b], d["rs"][b], d["c"], d["f"], itype=itype, engine="fp64")
        emax, _, mism = O.column_error(got[b], want)
        assert mism == 0 and emax <= TOL, (b, emax)


def test_batch_graph_replay(ccfj, batch):
    import torch
    d = batch
    ctx = ccfj._context(0)
    Us, Rs = [], []
    for uf, r in zip(d["ufs"], d["rs"]):
        idx = np.argsort(r)
        Us.append(np.asarray(uf)[idx].real.astype(np.float32))
        Rs.append(np.asarray(r)[idx].astype(np.float32))
    soff = np.zeros(len(Us) + 1, np.int64)
    soff[1:] = np.cumsum([len(r) for r in Rs])
    U = np.ascontiguousarray(np.concatenate(Us))
    R = np.concatenate(Rs)
    f32 = lambda a: np.asarray(a, np.float32)
    eager = ctx.transform_batched(ccfj.MODE_J, 1, True, soff, U, None, R, f32(d["f"]), f32(d["c"]))
    dev = torch.device("cuda:0")
    tU, tR, tf, tc = (torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                      for a in (U, R, f32(d["f"]), f32(d["c"])))
    out = torch.empty(eager.shape, dtype=torch.float32, device=dev)
    plan = ctx.batch_plan(ccfj.MODE_J, 1, True, soff, len(d["f"]), len(d["c"]))
    s = torch.cuda.Stream(device=dev)
    torch.cuda.synchronize()
    for rep in range(4):
        out.fill_(-7.0)
        torch.cuda.synchronize()
        plan.run(tU.data_ptr(), tU.data_ptr(), tR.data_ptr(), tf.data_ptr(), tc.data_ptr(),
                 out.data_ptr(), s.cuda_stream, True)
        s.synchronize()
        got = out.cpu().numpy()
        assert np.array_equal(np.isnan(got), np.isnan(eager))
        fin = np.isfinite(eager)
        assert np.max(np.abs(got[fin] - eager[fin])) <= 1e-6, rep
        _ = (tU * 2).sum().item()      # unrelated eager work between replays
    assert plan.tile in (64, 128, 256)
