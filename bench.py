#!/usr/bin/env python3
"""F-J throughput benchmark on one MI355X.

Workload (the dense noise-array configuration): 200 stations -> 19,900 pairs,
nf = 4096 frequencies (0.005-2 Hz), nc = 3000 velocities (500-6000 m/s),
three-mode synthetic cross-correlation spectra plus noise.  One step is
fj_noise for all four (itype, func) variants -- trapezoid / linear segments
x Bessel / Hankel basis -- on spectra already resident in HBM, each
normalised per frequency column on the device, as the API returns them.

For func = 1 the step includes the Hilbert transform of the spectra along
frequency (double-precision rFFT on the device), as fj_noise does.

Term-evals: nf * nc * nr per reference-equivalent native call; func = 0 is
one call, func = 1 two (FJ and FH), so a step is 6 calls = 1.467e12.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]

Prints one JSON line.  Besides the contract fields it carries
  roofline      the dominant kernel's algorithmic FP32 rate (34 flop per
                trapezoid term, 92 per linear-segment term and basis call)
                from HIP-event timing of that kernel, against 157.3 TFLOP/s
  cpu_baseline  the FJcpu restatement (oracle/, OpenMP, the host's cores)
                on a bounded sample of frequency columns
  plain_fp64_gpu  the plain single-instance FP64 GPU F-J on a column sample
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "cc-fjpy_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("F-J term-evals/s (f×c×r) at ≤1e-5 rel err vs FP64; "
          "speedup vs plain GPU F-J")
VARIANTS = [(0, 0), (0, 1), (1, 0), (1, 1)]   # (itype, func)
FLOP_PER_TERM = {0: 34, 1: 92}                # per basis call
FP32_PEAK = 157.3                             # TFLOP/s, MI355X
HBM_PEAK = 8000.0                             # GB/s
PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_fj_dense_noise_array_pmc.md")


def calls(func):
    return 1 if func == 0 else 2


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--variants", default="all", help="'all' or e.g. '1:1,0:0' (itype:func)")
    p.add_argument("--kernel-reps", type=int, default=2)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--copy-calib", action="store_true",
                   help="only run a device copy of known size (PMC byte-counter calibration)")
    return p.parse_args()


def measured_traffic(key):
    """HBM bytes per launch of the F-J kernel for variant key, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or
    None when that variant was not profiled."""
    try:
        with open(PMC_SUMMARY) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == key:
                    return int(parts[2])
    except OSError:
        pass
    return None


def copy_calibration():
    import torch
    n = 1 << 28   # 1 GiB of float32
    a = torch.ones(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 5}))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build measures one GPU. "
              "The F-J path shards by frequency with no communication (see DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        copy_calibration()
        return
    import torch
    import ccfj
    from ccfj import synth

    variants = VARIANTS if args.variants == "all" else [
        tuple(int(x) for x in v.split(":")) for v in args.variants.split(",")]
    print("bench: generating the dense noise-array workload", file=sys.stderr, flush=True)
    d = synth.noise_dense(real_only=True)
    idx = np.argsort(d["r"])
    U = np.ascontiguousarray(d["uf"][idx])
    r = d["r"][idx].astype(np.float32)
    f = d["f"].astype(np.float32)
    c = d["c"].astype(np.float32)
    nr, nf, nc = U.shape[0], U.shape[1], len(c)
    dev = torch.device("cuda:0")
    tU = torch.from_numpy(U).to(dev)
    tH = torch.empty_like(tU)
    tr, tf, tc = (torch.from_numpy(a).to(dev) for a in (r, f, c))
    tout = torch.empty((nc, nf), dtype=torch.float32, device=dev)
    ctx = ccfj._context(0)
    # a dedicated stream (default priority): the library launches on it and
    # the step events are recorded on it
    tstream = torch.cuda.Stream(device=dev)
    stream = tstream.cuda_stream
    torch.cuda.synchronize()

    def run(itype, func):
        if func == 1:
            ctx.hilbert_dev(tU.data_ptr(), tH.data_ptr(), nr, nf, stream)
        mode = ccfj.MODE_J if func == 0 else ccfj.MODE_J_MINUS_Y
        ctx.transform_dev(mode, itype, True, tU.data_ptr(), tH.data_ptr(), tr.data_ptr(),
                          tf.data_ptr(), tc.data_ptr(), nr, nf, nc, tout.data_ptr(), stream)

    def step():
        for it, fn in variants:
            run(it, fn)

    print("bench: warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    if not any(fn == 1 for _, fn in variants):
        ctx.hilbert_dev(tU.data_ptr(), tH.data_ptr(), nr, nf, stream)
        torch.cuda.synchronize()
    H = tH.cpu().numpy()    # the device Hilbert transform, for the baselines
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(tstream)
    for k in range(args.steps):
        step()
        print(f"bench: step {k + 1}/{args.steps} enqueued", file=sys.stderr, flush=True)
    e1.record(tstream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = gpu_ms / args.steps
    terms_step = sum(synth.term_evals(nr, nf, nc, calls(fn)) for _, fn in variants)
    value = terms_step * args.steps / (gpu_ms / 1e3)

    # per-kernel HIP-event timing of the main F-J kernels (library records
    # events on the launch stream around each main kernel)
    kernels = {}
    ctx.set_timing(True)
    for it, fn in variants:
        for _ in range(args.kernel_reps):
            run(it, fn)
        torch.cuda.synchronize()
        ms, n = ctx.get_timing()
        terms = synth.term_evals(nr, nf, nc, calls(fn))
        avg = ms / max(n, 1)
        kernels[f"itype{it}_func{fn}"] = {
            "avg_ms": round(avg, 4), "launches": n,
            "term_evals_per_launch": terms,
            "tflops_alg": round(terms * FLOP_PER_TERM[it] / (avg / 1e3) / 1e12, 3),
        }
    ctx.set_timing(False)
    dom_key = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    dom = kernels[dom_key]
    achieved = dom["tflops_alg"]
    roofline = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK, 4), "traffic": measured_traffic(dom_key),
                "kernel": f"k_fj ({dom_key})", "avg_ms": dom["avg_ms"]}
    print(f"bench: kernels {json.dumps(kernels)}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": value, "unit": "term-evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "dense noise array fj_noise, itype {0,1} x func {0,1}, "
                               "device-resident spectra, per-column normalisation",
                   "pairs": nr, "nf": nf, "nc": nc,
                   "variants": [f"itype{a}_func{b}" for a, b in variants],
                   "term_evals_per_step": terms_step},
        "roofline": roofline,
        "kernels": kernels,
        "wall_s_timed": round(wall, 3),
    }

    if not args.no_plain:
        from ccfj import _ccfj
        cols = np.arange(0, nf, 128)
        Us, Hs, fs = np.ascontiguousarray(U[:, cols]), np.ascontiguousarray(H[:, cols]), f[cols]
        pt, pms = 0.0, 0.0
        print("bench: plain FP64 GPU baseline", file=sys.stderr, flush=True)
        for it, fn in variants:
            _, ms = _ccfj.plain_fp64(0, 0, it, Us, r, fs, c)
            pms += ms
            pt += synth.term_evals(nr, len(cols), nc)
            if fn == 1:
                _, ms = _ccfj.plain_fp64(0, 1, it, Hs, r, fs, c)
                pms += ms
                pt += synth.term_evals(nr, len(cols), nc)
        prate = pt / (pms / 1e3)
        out["plain_fp64_gpu"] = {"value": prate, "unit": "term-evals/s",
                                 "sample": f"{len(cols)} of {nf} frequency columns (every 128th), "
                                           "all variants, kernel time"}
        out["speedup_vs_plain_fp64_gpu"] = round(value / prate, 2)

    if not args.no_cpu_baseline:
        from oracle import oracle as O
        cores = O.nthreads_default()
        cols = np.array([2048], np.int32)
        print(f"bench: CPU baseline on {cores} threads", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        terms = 0.0
        for it, fn in variants:
            O.port_columns(U, r, f, c, it, 0, cols, nthreads=cores)
            terms += synth.term_evals(nr, len(cols), nc)
            if fn == 1:
                O.port_columns(H, r, f, c, it, 1, cols, nthreads=cores)
                terms += synth.term_evals(nr, len(cols), nc)
        dt = time.perf_counter() - t0
        cpu_model = ""
        try:
            with open("/proc/cpuinfo") as fh:
                cpu_model = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
        except Exception:
            pass
        out["cpu_baseline"] = {"value": terms / dt, "unit": "term-evals/s", "cores": cores,
                               "kind": "port",
                               "sample": f"frequency columns {cols.tolist()} of {nf}, all {nc} "
                                         f"velocities and {nr} pairs, every variant "
                                         f"({terms:.3e} term-evals in {dt:.1f} s)",
                               "cpu": cpu_model}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
