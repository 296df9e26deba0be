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

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload dense]

Other workloads (--workload), each printing its own JSON line:
  batched   4096 sub-arrays of 20-400 pairs (nf 512, nc 1024, itype 1,
            func 0) in one persistent launch replayed from a HIP graph
  mwfj      multi-window F-J of 480 receivers x 16384 samples, 64 windows,
            nc 2000, f < 5 Hz, itype 0, func 0 (windows-as-RHS MFMA)
  learned   the band-proposal CNN: training on device-generated spectra,
            then banded dense F-J (nc 4000) against the full grid
  cc        the cross-correlation producer: 200 stations, one day at 4 Hz,
            one-hour windows at 90% overlap, 1000 bins of stride 3, whitened

Prints one JSON line.  Besides the contract fields it carries
  roofline      the dominant kernel's algorithmic FP32 rate (34 flop per
                trapezoid term, 92 per linear-segment term and basis call)
                from HIP-event timing of that kernel, against 157.3 TFLOP/s,
                and the FP32 rate it executes (committed SQ counter summary)
  cpu_baseline  the FJcpu restatement (oracle/, OpenMP, the host's cores)
                on a bounded sample of frequency columns; the same CPU leg
                checks the timed outputs against the FP64 oracle (accuracy)
  plain_fp64_gpu  the plain single-instance FP64 GPU F-J on 256 frequency
                columns (enough workgroups to fill the device)
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
FP16_MFMA_PEAK = 2500.0                       # TFLOP/s, MI355X dense fp16 matrix
HBM_PEAK = 8000.0                             # GB/s
PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_fj_dense_noise_array_pmc.md")
VALU_SUMMARY = os.path.join(ROOT, "profiles", "k_fj_dense_noise_array_valu.md")
CC_SUMMARY = os.path.join(ROOT, "profiles", "k_cc_gram_noise_day.md")
MWFJ_SUMMARY = os.path.join(ROOT, "profiles", "k_mwfj_mfma_active_source.md")
BATCH_SUMMARY = os.path.join(ROOT, "profiles", "k_fj_b_batched_subarrays.md")


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
    p.add_argument("--cpu-scale", type=int, default=1,
                   help="dense: frequency columns per thread in the CPU baseline sample "
                        "(1 takes about 25 s on 16 cores)")
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--no-host", action="store_true", help="dense: skip the host end-to-end call")
    p.add_argument("--workload", default="dense",
                   choices=["dense", "batched", "mwfj", "learned", "cc"])
    p.add_argument("--train", type=int, default=50000, help="learned: training spectra")
    p.add_argument("--epochs", type=int, default=3, help="learned: training epochs")
    p.add_argument("--copy-calib", action="store_true",
                   help="only run a device copy of known size (PMC byte-counter calibration)")
    p.add_argument("--valu-calib", action="store_true",
                   help="only run kernels of known instruction counts per class (issue cost "
                        "and SQ counter calibration)")
    return p.parse_args()


def summary_value(path, tag, key, conv):
    try:
        with open(path) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == tag and parts[1] == key:
                    return conv(parts[2])
    except OSError:
        pass
    return None


def measured_traffic(key):
    """HBM bytes per launch of the F-J kernel for variant key, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or
    None when that variant was not profiled."""
    return summary_value(PMC_SUMMARY, "traffic_bytes_per_launch", key, int)


def executed_tflops(key):
    """FP32 TFLOP/s the F-J kernel of variant key executes (2 FMA + ADD + MUL
    instructions x 64 lanes over its duration), from the committed rocprofv3
    SQ counter summary, or None."""
    return summary_value(VALU_SUMMARY, "executed_fp32_tflops", key, float)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            return next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
    except Exception:
        return ""


def host_info():
    """CPU model, the cores of this process's affinity mask and the host's
    OMP_NUM_THREADS, beside the threads a CPU baseline used ("cores")."""
    from oracle import oracle as O
    return {"cpu": cpu_model(), "host_cores": O.host_cores(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def copy_calibration():
    import torch
    n = 1 << 28   # 1 GiB of float32
    a = torch.ones(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 5}))


VALU_KINDS = ["v_fma_f32", "v_mul_f32", "v_add_f32", "v_sin_f32", "v_rcp_f32", "v_fma_f64",
              "v_fract_f64", "v_add_u32", "v_cndmask_b32", "v_cvt_f32_f64"]


def valu_calibration():
    """Issue cost per instruction class: kernels of 32 * iters instructions
    of one class per wave (cc-fjpy_amd/csrc/valu_calib.hip), timed with HIP
    events, at 1, 2, 4 and 8 waves per SIMD; the engine clock from
    s_memtime against the 100 MHz s_memrealtime counter.  Prints one JSON
    line: per class and occupancy, cycles per wave64 instruction per SIMD."""
    import ctypes
    import torch
    lib = ctypes.CDLL(os.path.join(ROOT, "cc-fjpy_amd", "ccfj", "libccfj_calib.so"))
    lib.calib_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                              ctypes.POINTER(ctypes.c_double)]
    lib.calib_clock.argtypes = [ctypes.c_void_p, ctypes.c_int]
    assert lib.calib_kinds() == len(VALU_KINDS)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    simds = 4 * cus
    ck = torch.zeros(3, dtype=torch.int64, device="cuda:0")
    spins = 2_000_000
    assert lib.calib_clock(ck.data_ptr(), spins) == 0
    ticks, real = (int(v) for v in ck.cpu().tolist()[:2])
    clock_ghz = ticks / (real / 100e6) / 1e9
    iters = 4000
    res = {"cus": cus, "simds": simds, "clock_ghz_s_memtime": round(clock_ghz, 4),
           "instructions_per_wave": 32 * iters, "classes": {}}
    for wps in (1, 2, 4, 8):
        nblocks = simds * wps // 4          # 256 lanes = 4 waves per workgroup
        out = torch.empty(nblocks * 256, dtype=torch.float32, device="cuda:0")
        waves = nblocks * 4
        for k, name in enumerate(VALU_KINDS):
            ms = ctypes.c_double()
            for _ in range(2):              # the second run is the one kept
                assert lib.calib_run(k, nblocks, iters, out.data_ptr(), ctypes.byref(ms)) == 0
            per_simd = waves * 32.0 * iters / simds
            cyc = ms.value * 1e-3 * clock_ghz * 1e9 / per_simd
            res["classes"].setdefault(name, {})[f"{wps}_waves_per_simd"] = {
                "ms": round(ms.value, 4), "cycles_per_instruction_per_simd": round(cyc, 3)}
        print(f"bench: valu calibration at {wps} waves/SIMD done", file=sys.stderr, flush=True)
    print(json.dumps(res), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build measures one GPU. "
              "The F-J path would shard by frequency columns, with no communication.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        copy_calibration()
        return
    if args.valu_calib:
        valu_calibration()
        return
    if args.workload != "dense":
        {"batched": bench_batched, "mwfj": bench_mwfj, "learned": bench_learned,
         "cc": bench_cc}[args.workload](args)
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
    # one result array per variant, so that the timed outputs stay for the
    # accuracy check after the timed region
    touts = {v: torch.empty((nc, nf), dtype=torch.float32, device=dev) for v in variants}
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
                          tf.data_ptr(), tc.data_ptr(), nr, nf, nc,
                          touts[(itype, func)].data_ptr(), stream)

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
    # the last timed step's outputs at the accuracy columns (0.49 and 2 Hz)
    acc_cols = [1365, nf - 1]
    timed_cols = {v: touts[v][:, acc_cols].cpu().numpy() for v in variants}
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
    ex = executed_tflops(dom_key)
    roofline = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK, 4), "traffic": measured_traffic(dom_key),
                "kernel": f"k_fj ({dom_key})", "avg_ms": dom["avg_ms"],
                "achieved_is": "algorithmic flops (92 per linear-segment term and basis call) per "
                               "launch / HIP-event launch time; a frac above 1 means the kernel "
                               "needs fewer flops per term than that closed-form recipe (the "
                               "station-only work is precomputed into the records): "
                               "executed_frac is the share of the FP32 pipe in use",
                "executed_fp32_tflops": ex,
                "executed_frac": None if ex is None else round(ex / FP32_PEAK, 4)}
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

    if not args.no_host and (1, 0) in variants:
        # the reference call end to end: complex64 spectra in host memory,
        # unsorted distances, numpy float32 [nc, nf] out (sort, cast, H2D,
        # transform, normalise, D2H)
        import warnings
        ufc = d["uf"].astype(np.complex64)
        print("bench: fj_noise end to end from host memory", file=sys.stderr, flush=True)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            ccfj.fj_noise(ufc, d["r"], d["c"], d["f"], itype=1, func=0, num=0)
            t0 = time.perf_counter()
            ccfj.fj_noise(ufc, d["r"], d["c"], d["f"], itype=1, func=0, num=0)
            dt = time.perf_counter() - t0
        del ufc
        out["end_to_end_host"] = {
            "value": synth.term_evals(nr, nf, nc) / dt, "unit": "term-evals/s",
            "ms_per_call": round(dt * 1e3, 1),
            "what": "fj_noise(uf, r, c, f, itype=1, func=0): complex64 spectra in host memory, "
                    "unsorted r, numpy float32 [nc, nf] out"}

    if not args.no_plain:
        from ccfj import _ccfj
        # every 16th column: 256 frequencies x 3000 velocities in 32 x 16
        # blocks is 8 x 188 = 1504 workgroups, enough to fill the 256 CUs
        cols = np.arange(0, nf, 16)
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
                                 "sample": f"{len(cols)} of {nf} frequency columns (every 16th), "
                                           f"all {nc} velocities and {nr} pairs, all variants, "
                                           "kernel time",
                                 "workgroups": ((len(cols) + 31) // 32) * ((nc + 15) // 16)}
        out["speedup_vs_plain_fp64_gpu"] = round(value / prate, 2)

    if not args.no_cpu_baseline:
        from oracle import oracle as O
        cores = O.nthreads_default()
        # the restatement parallelises over frequency columns (OpenMP over f,
        # as FJcpu): one column per thread, every 4th velocity
        cols = np.linspace(128, nf - 1, cores * max(1, args.cpu_scale)).astype(np.int32)
        cs = np.ascontiguousarray(c[::4])
        print(f"bench: CPU baseline on {cores} threads", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        terms = 0.0
        for it, fn in variants:
            O.port_columns(U, r, f, cs, it, 0, cols, nthreads=cores)
            terms += synth.term_evals(nr, len(cols), len(cs))
            if fn == 1:
                O.port_columns(H, r, f, cs, it, 1, cols, nthreads=cores)
                terms += synth.term_evals(nr, len(cols), len(cs))
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": terms / dt, "unit": "term-evals/s", "cores": cores,
                               "kind": "port",
                               "sample": f"{len(cols)} frequency columns of {nf} "
                                         f"({len(cols) // cores} per thread), "
                                         f"every 4th of {nc} velocities, all {nr} pairs, every "
                                         f"variant ({terms:.3e} term-evals in {dt:.1f} s)",
                               **host_info()}
        # the same CPU leg checks the timed outputs: e = |X - X*| / max_c |X*[:, f]|
        # against the FP64 oracle on the same float32 inputs, with the Hilbert
        # transform of the oracle side taken on the host (numpy, float64)
        print("bench: accuracy of the timed outputs against the FP64 oracle", file=sys.stderr,
              flush=True)
        Hh = ccfj.hilbert_rows(U)
        ref = {}
        acc = {"columns": acc_cols, "f_hz": [float(f[i]) for i in acc_cols], "bar": 1e-5,
               "reference": "FP64 oracle (oracle/fj_oracle64.c) on the float32 inputs, "
                            "per-column normalised"}
        for it, fn in variants:
            if (it, 0) not in ref:
                ref[it, 0] = O.oracle64(U, r, f, c, it, 0, fidx=acc_cols, nthreads=cores)
            want = ref[it, 0]
            if fn == 1:
                if (it, 1) not in ref:
                    ref[it, 1] = O.oracle64(Hh, r, f, c, it, 1, fidx=acc_cols, nthreads=cores)
                want = want - ref[it, 1]
            want = want / np.max(np.abs(want), axis=0)
            emax, erms, mism = O.column_error(timed_cols[(it, fn)], want)
            acc[f"itype{it}_func{fn}"] = {"e_max": float(f"{emax:.3e}"), "e_rms": float(f"{erms:.3e}"),
                                          "nonfinite_mismatch": mism}
        acc["all_within_bar"] = all(acc[f"itype{a}_func{b}"]["e_max"] <= 1e-5 and
                                    acc[f"itype{a}_func{b}"]["nonfinite_mismatch"] == 0
                                    for a, b in variants)
        out["accuracy"] = acc
    print(json.dumps(out), flush=True)


def contract(metric, value, unit, args, ms_per_step, config, dtype="f32"):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
            "config": config}


def bench_batched(args):
    """4096 independent sub-arrays in one persistent launch (HIP graph)."""
    import torch
    import ccfj
    from ccfj import _ccfj, synth
    print("bench: generating 4096 sub-arrays", file=sys.stderr, flush=True)
    d = synth.subarrays(seed=3)
    ns = np.array([len(r) for r in d["rs"]])
    soff = np.zeros(len(ns) + 1, np.int64)
    soff[1:] = np.cumsum(ns)
    U = np.concatenate(d["ufs"], axis=0)
    r = np.concatenate(d["rs"]).astype(np.float32)
    f, c = d["f"].astype(np.float32), d["c"].astype(np.float32)
    nb, nf, nc = len(ns), len(f), len(c)
    dev = torch.device("cuda:0")
    tU, tr, tf, tc = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (U, r, f, c))
    tout = torch.empty((nb, nc, nf), dtype=torch.float32, device=dev)
    ctx = ccfj._context(0)
    tstream = torch.cuda.Stream(device=dev)
    stream = tstream.cuda_stream
    plan = ctx.batch_plan(ccfj.MODE_J, 1, True, soff, nf, nc)
    torch.cuda.synchronize()

    def run():
        plan.run(tU.data_ptr(), tU.data_ptr(), tr.data_ptr(), tf.data_ptr(), tc.data_ptr(),
                 tout.data_ptr(), stream, True)

    print("bench: tuning + graph capture + warmup", file=sys.stderr, flush=True)
    run()
    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(tstream)
    for k in range(args.steps):
        run()
    e1.record(tstream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    terms = float(ns.sum()) * nf * nc
    value = terms / (ms / 1e3)
    achieved = terms * FLOP_PER_TERM[1] / (ms / 1e3) / 1e12
    out = contract(METRIC, value, "term-evals/s", args, ms,
                   {"workload": "batched sub-arrays fj_noise itype 1 func 0, one persistent "
                                "launch replayed from a HIP graph, device-resident spectra",
                    "sub_arrays": nb, "pairs_total": int(ns.sum()), "nf": nf, "nc": nc,
                    "tile": plan.tile, "term_evals_per_step": terms})
    out["roofline"] = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK,
                       "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK, 4),
                       "traffic": summary_value(BATCH_SUMMARY, "traffic_bytes_per_launch",
                                                "k_fj_b", int),
                       "kernel": "whole graph (stations, prep, persistent k_fj_b, finish)",
                       "achieved_is": "algorithmic flops (92 per linear-segment term) / the "
                                      "graph's time; traffic is the persistent kernel's",
                       "executed_fp32_tflops": summary_value(BATCH_SUMMARY, "executed_fp32_tflops",
                                                             "k_fj_b", float)}
    if not args.no_plain:
        print("bench: plain FP64 GPU baseline (every 64th sub-array)", file=sys.stderr, flush=True)
        pt, pms = 0.0, 0.0
        for b in range(0, nb, 64):
            Ub = np.ascontiguousarray(U[soff[b]:soff[b + 1]])
            _, kms = _ccfj.plain_fp64(0, 0, 1, Ub, r[soff[b]:soff[b + 1]], f, c)
            pms += kms
            pt += synth.term_evals(ns[b], nf, nc)
        prate = pt / (pms / 1e3)
        out["plain_fp64_gpu"] = {"value": prate, "unit": "term-evals/s",
                                 "sample": f"{len(range(0, nb, 64))} of {nb} sub-arrays "
                                           "(every 64th), one call each, kernel time"}
        out["speedup_vs_plain_fp64_gpu"] = round(value / prate, 2)
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        cores = O.nthreads_default()
        t0 = time.perf_counter()
        ct = 0.0
        picks = list(range(0, nb, nb // 4))[:4]
        for b in picks:
            Ub = np.ascontiguousarray(U[soff[b]:soff[b + 1]])
            O.port_native(Ub, r[soff[b]:soff[b + 1]], f, c, 1, 0, nthreads=cores)
            ct += synth.term_evals(ns[b], nf, nc)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": ct / dt, "unit": "term-evals/s", "cores": cores,
                               "kind": "port",
                               "sample": f"sub-arrays {picks} in full ({ct:.3e} term-evals in "
                                         f"{dt:.1f} s)", **host_info()}
    print(json.dumps(out), flush=True)


def bench_mwfj(args):
    """MWFJ on device-resident traces (ccfj_mwfj_dev: windowing, rFFT, MFMA
    contraction, normalisation), its MFMA contraction alone, and the host
    entry point end to end (numpy in, numpy out)."""
    import torch
    import ccfj
    from ccfj import _ccfj, synth
    d = synth.active_source(seed=2)
    u, r, f, c = d["u"], d["r"], d["f"], d["c"]
    nwin, nr, nf, nc = d["nwin"], len(r), len(f), len(c)
    npts = u.shape[1]
    ctx = ccfj._context(0)
    dev = torch.device("cuda:0")
    indx = np.argsort(r)
    n1, n2, dl, xidx, xrows = ccfj._window_descriptions(
        np.asarray(d["winl"])[:nwin], np.asarray(d["winr"])[:nwin], indx, npts, d["Fs"], 0.9)
    host = (np.ascontiguousarray(u[indx], np.float64), r[indx].astype(np.float32),
            f.astype(np.float32), c.astype(np.float32), n1, n2, dl, xidx,
            xrows if len(xrows) else np.zeros((1, npts)))
    dv = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in host]
    dout = torch.empty((nwin, nc, nf), dtype=torch.float64, device=dev)
    st = torch.cuda.Stream(device=dev)

    def run_dev():
        ctx.mwfj_dev(0, 0, 1, nr, npts, nwin, nf, nc, *[a.data_ptr() for a in dv], len(xrows),
                     dout.data_ptr(), st.cuda_stream)

    def run_host():
        return ccfj.MWFJ(u, r, c, f, d["Fs"], nwin, d["winl"], d["winr"], 0.9, 1, 0, 0, 0)

    print("bench: MWFJ warmup", file=sys.stderr, flush=True)
    for _ in range(max(1, args.warmup)):
        run_dev()
        out0 = run_host()
    st.synchronize()
    assert np.array_equal(dout.cpu().numpy(), out0)
    ctx.set_timing(True)
    ctx.get_timing()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_dev()
    st.synchronize()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    kms, nk = ctx.get_timing()
    ctx.set_timing(False)
    kms /= max(args.steps, 1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_host()
    wall_host = (time.perf_counter() - t0) / args.steps
    terms = synth.term_evals(nr, nf, nc, 2 * nwin)   # fj_earthquake func 0: Re and Im calls
    # contraction flops: 2 per (f, c, r, right-hand side), 2 nwin right-hand sides
    mfma_tflops = 2.0 * nf * nc * nr * 2 * nwin / (kms / 1e3) / 1e12
    out = contract(METRIC, terms / wall, "term-evals/s", args, wall * 1e3,
                   {"workload": "MWFJ on device-resident traces (float64 [nwin, nc, nf] out in "
                                "HBM), itype 0 func 0, windows as right-hand sides of one MFMA "
                                "contraction per frequency",
                    "receivers": nr, "npts": npts, "nwin": nwin, "nf": nf, "nc": nc,
                    "term_evals_per_step": terms})
    out["end_to_end_host"] = {"value": terms / wall_host, "unit": "term-evals/s",
                              "ms_per_step": round(wall_host * 1e3, 3),
                              "what": "MWFJ() with numpy in and out: H2D of the traces, the "
                                      "same device work, D2H of the float64 result"}
    out["contraction"] = {"kernel_ms": round(kms, 3), "launches_per_step": nk / max(args.steps, 1),
                          "term_evals_per_s": terms / (kms / 1e3),
                          "fp32_equivalent_tflops": round(mfma_tflops, 2)}
    # the products run as three fp16 products each on the fp16 matrix pipe
    # (f > 0): priced against its dense peak; the executed fp16 rate comes
    # from the committed counter summary
    avg_ms = kms / max(nk / max(args.steps, 1), 1)
    exe = summary_value(MWFJ_SUMMARY, "executed_flops_per_launch", "k_mwfj_mfma", float)
    out["roofline"] = {"bound": "mfma", "achieved": round(mfma_tflops, 2), "peak": FP16_MFMA_PEAK,
                       "unit": "TFLOP/s", "frac": round(mfma_tflops / FP16_MFMA_PEAK, 4),
                       "traffic": summary_value(MWFJ_SUMMARY, "traffic_bytes_per_launch",
                                                "k_mwfj_mfma", int),
                       "kernel": "k_mwfj_mfma", "avg_ms": round(avg_ms, 4),
                       "achieved_is": "algorithmic FP32 contraction flops (2 per f, c, station "
                                      "and right-hand side) / HIP-event time of the contraction",
                       "peak_is": "dense fp16 MFMA (v_mfma_f32_32x32x16_f16: three split products "
                                  "per FP32 product at f > 0)",
                       "executed_fp16_tflops": round(exe / (avg_ms / 1e3) / 1e12, 1) if exe else None,
                       "executed_frac": round(exe / (avg_ms / 1e3) / 1e12 / FP16_MFMA_PEAK, 4)
                       if exe else None,
                       "mfma_busy": summary_value(MWFJ_SUMMARY, "mfma_busy", "k_mwfj_mfma", float)}
    if not args.no_plain:
        print("bench: plain FP64 GPU baseline (4 windows)", file=sys.stderr, flush=True)
        idx = np.argsort(r)
        pt, pms = 0.0, 0.0
        for w in range(0, nwin, nwin // 4):
            tap = np.stack([ccfj.win(u.shape[1], d["Fs"], d["winl"][w, j], d["winr"][w, j], 0.9)
                            for j in range(nr)])
            spec = np.fft.rfft(u.astype(np.float64) * tap, axis=1)[idx][:, :nf]
            for part in (spec.real, spec.imag):
                _, kms2 = _ccfj.plain_fp64(0, 0, 0, np.ascontiguousarray(part, np.float32),
                                           r[idx].astype(np.float32), f.astype(np.float32),
                                           c.astype(np.float32))
                pms += kms2
                pt += synth.term_evals(nr, nf, nc)
        prate = pt / (pms / 1e3)
        out["plain_fp64_gpu"] = {"value": prate, "unit": "term-evals/s",
                                 "sample": "4 of 64 windows, Re and Im calls, kernel time"}
        out["speedup_vs_plain_fp64_gpu"] = round(out["value"] / prate, 2)
    if not args.no_cpu_baseline:
        # the FJcpu restatement (itype 0, J basis) on the real parts of the
        # first 4 windows: 4 of the 2 nwin calls per step (OpenMP over f, as
        # FJcpu)
        from oracle import oracle as O
        cores = O.nthreads_default()
        idx = np.argsort(r)
        nw = 4
        dt, ct = 0.0, 0.0
        for w in range(nw):
            tap = np.stack([ccfj.win(u.shape[1], d["Fs"], d["winl"][w, j], d["winr"][w, j], 0.9)
                            for j in range(nr)])
            spec = np.fft.rfft(u.astype(np.float64) * tap, axis=1)[idx][:, :nf]
            t0 = time.perf_counter()
            O.port_native(np.ascontiguousarray(spec.real, np.float32), r[idx], f, c, 0, 0,
                          nthreads=cores)
            dt += time.perf_counter() - t0
            ct += synth.term_evals(nr, nf, nc)
        out["cpu_baseline"] = {"value": ct / dt, "unit": "term-evals/s", "cores": cores,
                               "kind": "port",
                               "sample": f"{nw} FJ calls (real parts of windows 0-{nw - 1}), all "
                                         f"{nf} frequencies and {nc} velocities ({ct:.3e} "
                                         f"term-evals in {dt:.1f} s; F-J only, no windowing or "
                                         f"FFT)",
                               **host_info()}
    print(json.dumps(out), flush=True)


def bench_learned(args):
    """Band-proposal CNN: train on device-generated coarse spectra, then the
    banded dense F-J against the full grid."""
    import torch  # noqa: F401
    from ccfj import learn
    m = learn.Coarse2Fine(num=0, seed=0)
    t0 = time.perf_counter()
    X, Y, _ = m.make_set(args.train, seed=1)
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = m.train(X, Y, args.train, epochs=args.epochs, lr=2e-3,
                     log=lambda s: print("bench:", s, file=sys.stderr, flush=True))
    m.ctx.synchronize()
    t_train = time.perf_counter() - t0
    res = m.evaluate(64, seed=7, levels=(0.5, 0.3, 0.2))
    cnn = res["cnn"]
    full_ms = res["full_kernel_ms"]
    # the banded path's cost: coarse F-J + network + banded dense kernel
    band_ms = res["coarse_ms"] + res["cnn_ms"] + cnn["kernel_ms"]
    dense_terms = 64 * m.nr * m.W * m.nc_dense
    value = dense_terms / (band_ms / 1e3)   # full-grid-equivalent term-evals/s of the banded path
    out = contract("banded dense F-J, full-grid-equivalent term-evals/s at the ridge agreement "
                   "below", value, "term-evals/s", args, band_ms / 64,
                   {"workload": "learned coarse-to-fine: coarse 64c x 256f F-J of 300-pair fields, "
                                "CNN bands, dense nc 4000 (0.01-1 Hz) only inside them",
                    "training_spectra": args.train, "epochs": args.epochs, "eval_fields": 64})
    out["ridge"] = res
    out["training"] = {"data_s": round(t_data, 2), "train_s": round(t_train, 2),
                       "steps": len(losses), "first_loss": float(losses[0]),
                       "last_loss": float(np.mean(losses[-20:]))}
    out["speedup_vs_full_grid"] = round(full_ms / band_ms, 2)
    if not args.no_cpu_baseline:
        # a CPU has no bands: the FJcpu restatement (OpenMP over f) on the
        # full dense grid of one field of the same kind
        from ccfj import learn as L
        from oracle import oracle as O
        cores = O.nthreads_default()
        rng = np.random.default_rng(123)
        r, modes, sigma = L.sample_fields(rng, 1, m.nr)
        U, d_r, _ = m._fields(r, modes, sigma, 99)
        m.ctx.synchronize()
        Uh = U.download().reshape(m.nr, m.W)
        t0 = time.perf_counter()
        O.port_native(Uh, np.asarray(r, np.float32).reshape(-1), m.f, m.c_dense, m.itype, 0,
                      nthreads=cores)
        dt = time.perf_counter() - t0
        ct = float(m.nr) * m.W * m.nc_dense
        out["cpu_baseline"] = {"value": ct / dt, "unit": "term-evals/s", "cores": cores,
                               "kind": "port",
                               "sample": f"one field's full dense grid ({m.nr} pairs x {m.W} f x "
                                         f"{m.nc_dense} c, itype {m.itype}: {ct:.3e} term-evals "
                                         f"in {dt:.1f} s)", **host_info()}
    print(json.dumps(out), flush=True)


def bench_cc(args):
    """Cross-correlation of 200 stations over one day (the reference's
    example settings: 4 Hz, one-hour windows, 90% overlap, nf 1000,
    fstride 3, whitening), synthetic records with a shared random wavefield
    plus station noise; station 7 starts late and station 11 ends early."""
    import torch
    import ccfj
    rng = np.random.default_rng(5)
    Fs, nsta, nf, fstride = 4, 200, 1000, 3
    npts = 24 * 3600 * Fs
    fftlen = 3600 * Fs
    common = rng.normal(size=npts + 4000).astype(np.float32)
    lags = rng.integers(0, 4000, nsta)
    data = np.stack([common[l:l + npts] for l in lags]) + rng.normal(size=(nsta, npts)).astype(np.float32)
    data = np.ascontiguousarray(data.astype(np.float32).reshape(-1))
    se = np.tile(np.array([0, npts], np.int32), nsta)
    se[2 * 7] = npts // 4
    se[2 * 11 + 1] = 3 * npts // 4
    pairs = ccfj.GetStationPairs(nsta)
    npairs = len(pairs) // 2
    steplen = int(fftlen * (1 - 0.9))
    nshifts = 1 + (npts - fftlen) // steplen
    dev = torch.device("cuda:0")
    ctx = ccfj._context(0)
    tp = torch.from_numpy(np.ascontiguousarray(pairs, np.int32)).to(dev)
    ts = torch.from_numpy(np.ascontiguousarray(se, np.int32)).to(dev)
    tx = torch.from_numpy(data).to(dev)
    tre = torch.empty((npairs, nf), dtype=torch.float32, device=dev)
    tim = torch.empty((npairs, nf), dtype=torch.float32, device=dev)
    st = torch.cuda.Stream(device=dev)

    def run_dev():
        ctx.cross_correlate_dev(nsta, npts, nf, fstride, fftlen, steplen, False, True, npairs,
                                tp.data_ptr(), ts.data_ptr(), tx.data_ptr(), tre.data_ptr(),
                                tim.data_ptr(), st.cuda_stream)

    print("bench: CC warmup", file=sys.stderr, flush=True)
    for _ in range(max(1, args.warmup)):
        run_dev()
        out = ccfj.CC(npts, nsta, nf, fftlen, pairs, se, data, overlaprate=0.9, fstride=fstride)
    st.synchronize()
    assert np.array_equal(tre.cpu().numpy(), out.real) and np.array_equal(tim.cpu().numpy(), out.imag)
    ctx.set_timing(True)
    ctx.get_timing()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_dev()
    st.synchronize()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    gms, ng = ctx.get_timing()     # the Gram pair kernel (HIP events on the launch stream)
    ctx.set_timing(False)
    gms /= max(args.steps, 1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = ccfj.CC(npts, nsta, nf, fftlen, pairs, se, data, overlaprate=0.9, fstride=fstride)
    wall_host = (time.perf_counter() - t0) / args.steps
    # pair-window products actually formed: window k of station i is used
    # when k*steplen >= start_i and k*steplen + fftlen < end_i
    t_k = np.arange(nshifts)[:, None] * steplen
    use = (t_k >= se[0::2][None, :]) & (t_k + fftlen < se[1::2][None, :])
    pw = int(np.sum(use[:, pairs[0::2]] & use[:, pairs[1::2]]))
    res = contract("cross-correlation pair-window spectra/s (nf bins each)", pw / wall,
                   "pair-windows/s", args, wall * 1e3,
                   {"workload": "ambient-noise cross-correlation, device-resident records, "
                                "float32 re/im [npairs, nf] out in HBM", "stations": nsta,
                    "npts": npts, "fftlen": fftlen, "windows": nshifts, "pairs": npairs,
                    "nf": nf, "fstride": fstride, "pair_windows": pw})
    # dominant kernel: the Gram pair stage.  Per launch it reads every
    # station's spectra of the chunk once (8 bytes a value) and writes the
    # pair sums (8 bytes a pair and bin; a later chunk also reads them);
    # 8 flops (one complex multiply-add) per (pair, shared window, bin) on
    # the fp16 matrix cores (k_cc_gram_split: three split products per FP32
    # product), 70 flops a byte, under the 312 of the fp16 ridge
    # (2500 TFLOP/s / 8 TB/s): HBM is the bounding roof.
    launches = ng / max(args.steps, 1)
    sums = 8.0 * npairs * nf * (2.0 * launches - 1.0) / launches
    gram_bytes = 8.0 * nsta * nshifts * nf / launches + sums
    lms = gms / max(launches, 1.0)   # average launch
    gram_gbs = gram_bytes / (lms / 1e3) / 1e9 if lms > 0 else 0.0
    gram_tflops = 8.0 * pw * nf / launches / (lms / 1e3) / 1e12 if lms > 0 else 0.0
    res["roofline"] = {"bound": "hbm", "achieved": round(gram_gbs, 1), "peak": HBM_PEAK,
                       "unit": "GB/s", "frac": round(gram_gbs / HBM_PEAK, 4),
                       "traffic": summary_value(CC_SUMMARY, "traffic_bytes_per_launch",
                                                "cc_gram", int),
                       "kernel": "k_cc_gram_split", "avg_ms": round(lms, 4),
                       "launches_per_step": launches,
                       "algorithmic_bytes_per_launch": int(gram_bytes),
                       "achieved_is": "spectra read once + pair sums written, per launch / "
                                      "HIP-event time",
                       "mfma_tflops_algorithmic": round(gram_tflops, 2),
                       "mfma_peak_fp16": FP16_MFMA_PEAK,
                       "executed_mfma_tflops": summary_value(CC_SUMMARY, "executed_mfma_tflops",
                                                             "cc_gram", float)}
    res["end_to_end_host"] = {"value": pw / wall_host, "unit": "pair-windows/s",
                              "ms_per_step": round(wall_host * 1e3, 3),
                              "what": "CC() with numpy in and out: H2D of the records, the same "
                                      "device work, D2H of the complex64 result"}
    res["output_shape"] = list(out.shape)
    if not args.no_cpu_baseline:
        # the restatement on the whole job: every station, pair and window,
        # FFTs and pair sums on the host's threads
        from oracle import cc_ref as C
        from oracle import oracle as O
        cores = O.nthreads_default()
        print(f"bench: CPU baseline on {cores} threads", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        _, cnt = C.cross_correlate_threaded(data, nsta, npts, nf, fftlen, steplen, pairs, se,
                                            fstride=fstride, nthreads=cores)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": float(cnt.sum()) / dt, "unit": "pair-windows/s",
                               "cores": cores, "kind": "port",
                               "sample": f"threaded restatement (scipy float64 FFT) on all {nsta} "
                                         f"stations, {npairs} pairs and {nshifts} windows "
                                         f"({int(cnt.sum())} pair-windows in {dt:.1f} s)",
                               **host_info()}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
