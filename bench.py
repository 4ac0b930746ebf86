"""Benchmark: 512^3 ComplexF64 3D C2C, forward + inverse, Pencil, one GPU.

Methodology of the reference benchmark (benchmarks/fftmpigpu.jl:76-105):
warm-up, then timed iterations of fft! followed by ifft!;
time_per_transform = (t_fft + t_ifft) / 2 and GFLOP/s = 5 N log2(N) 1e-9 / t.

One "step" is one forward plus one inverse transform of the resident
512x512x512 ComplexF64 block (inputs generated on device by the SplitMix64
counter generator, seed 3).  Every axis pass is bracketed by HIP events on the
stream it runs on, giving per-pass GB/s against the 8 TB/s HBM peak.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "daggerfft.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--decomp", choices=["pencil", "slab"], default="pencil")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--profile-steps", action="store_true", help="no per-pass events (for rocprof runs)")
    return ap.parse_args()


def cpu_baseline(n, fwd_ms_gpu=None):
    """The oracle's restatement of the reference pipeline (per-axis batched
    1D FFTs in stage order) on the host cores, one full n^3 forward + inverse."""
    import numpy as np

    from oracle import oracle as O
    shape = (n, n, n)
    a = O.fill(3, shape, np.complex128)
    t0 = time.perf_counter()
    for ax in range(3):
        O.c2c_axis(a, ax, -1, 1.0)
    t1 = time.perf_counter()
    for ax in (2, 1, 0):
        O.c2c_axis(a, ax, +1, 1.0 / n)
    t2 = time.perf_counter()
    N = float(n) ** 3
    t = ((t1 - t0) + (t2 - t1)) / 2
    return {"value": round(5 * N * math.log2(N) * 1e-9 / t, 3), "unit": "GFLOP/s", "cores": O.num_threads(),
            "kind": "port", "sample": "one full %d^3 ComplexF64 forward + inverse through the C oracle "
            "(per-axis batched FFTs in the reference's stage order), time per transform %.3f s" % (n, n, t)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import daggerfft as D
    from daggerfft import _lib as L

    torch.cuda.set_device(0)
    n = args.n
    shape = (n, n, n)
    N = float(n) ** 3
    decomp = D.Pencil() if args.decomp == "pencil" else D.Slab()
    A = D.rand(shape, np.complex128, seed=3)
    X = A.similar()
    Y = A.similar()
    pf = D.Plan(shape, L.C2C, -1, L.F64, decomp, (1, 2, 3))
    pi = D.Plan(shape, L.C2C, +1, L.F64, decomp, (1, 2, 3))
    fpasses, ipasses = pf.passes(), pi.passes()

    def step(events=None):
        # forward A -> X, inverse X -> Y
        for i in range(len(fpasses)):
            if events is not None:
                events[("f", i)][0].record()
            pf.execute_pass(i, A, X)
            if events is not None:
                events[("f", i)][1].record()
        for i in range(len(ipasses)):
            if events is not None:
                events[("i", i)][0].record()
            pi.execute_pass(i, X, Y)
            if events is not None:
                events[("i", i)][1].record()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    keys = [("f", i) for i in range(len(fpasses))] + [("i", i) for i in range(len(ipasses))]
    per = {k: 0.0 for k in keys}
    t_fwd = t_inv = 0.0
    s0 = torch.cuda.Event(enable_timing=True)
    s1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s0.record()
    wall0 = time.perf_counter()
    ev_all = []
    for _ in range(args.steps):
        ev = None if args.profile_steps else {k: (torch.cuda.Event(enable_timing=True),
                                                  torch.cuda.Event(enable_timing=True)) for k in keys}
        step(ev)
        ev_all.append(ev)
    s1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = s0.elapsed_time(s1)
    if not args.profile_steps:
        for ev in ev_all:
            for k in keys:
                dt = ev[k][0].elapsed_time(ev[k][1])
                per[k] += dt
                if k[0] == "f":
                    t_fwd += dt
                else:
                    t_inv += dt
    ms_step = total_ms / args.steps
    t_transform_s = ms_step / 2 / 1e3
    gflops = 5 * N * math.log2(N) * 1e-9 / t_transform_s

    # round-trip check of the last step
    err = float(torch.linalg.vector_norm(Y.tensor - A.tensor) / torch.linalg.vector_norm(A.tensor))

    passes = []
    for k in keys:
        desc = (fpasses if k[0] == "f" else ipasses)[k[1]]
        avg = per[k] / args.steps if not args.profile_steps else float("nan")
        gbps = desc["bytes"] / (avg / 1e3) / 1e9 if avg and avg == avg else None
        passes.append({"dir": "fwd" if k[0] == "f" else "inv", "axis": desc["axis"] + 1, "op": desc["op"],
                       "L": desc["L"], "radices": desc["radices"], "TPL": desc["TPL"], "TC": desc["TC"],
                       "avg_ms": round(avg, 4), "GBps": round(gbps, 1) if gbps else None,
                       "frac": round(gbps / HBM_PEAK_GBPS, 4) if gbps else None})
    dom = max(passes, key=lambda p: p["avg_ms"]) if not args.profile_steps else None
    alg_bytes_transform = pf.algorithmic_bytes()

    # measured device-to-device copy ceiling on the same buffers
    torch.cuda.synchronize()
    c0 = torch.cuda.Event(enable_timing=True)
    c1 = torch.cuda.Event(enable_timing=True)
    X.tensor.copy_(A.tensor)
    c0.record()
    for _ in range(5):
        X.tensor.copy_(A.tensor)
    c1.record()
    torch.cuda.synchronize()
    copy_gbps = 2 * A.nbytes * 5 / (c0.elapsed_time(c1) / 1e3) / 1e9

    roofline = None
    if dom is not None:
        desc_bytes = (fpasses if dom["dir"] == "fwd" else ipasses)[0]["bytes"]
        roofline = {"bound": "hbm", "achieved": dom["GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": dom["frac"], "traffic": None,
                    "kernel": "k_axis<double,C2C,%s> axis %d (%s)" % ("contig" if dom["axis"] == 1 else "strided",
                                                                       dom["axis"], dom["dir"]),
                    "bytes_per_launch": desc_bytes, "avg_ms": dom["avg_ms"]}

    out = {
        "metric": "512^3 C2C FP64 GFLOP/s (5N*log2N) & %HBM peak; served-stream p50/p99 latency",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: SplitMix64 counter generator, seed 3, uniform[0,1) re/im, resident in HBM",
        "config": {"workload": "3D C2C ComplexF64 %dx%dx%d forward+inverse" % (n, n, n),
                   "decomp": args.decomp, "dims": [1, 2, 3], "step": "fft then ifft",
                   "time_per_transform_ms": round(ms_step / 2, 4),
                   "hbm_frac_per_transform": round(alg_bytes_transform / t_transform_s / 1e9 / HBM_PEAK_GBPS, 4),
                   "algorithmic_bytes_per_transform": alg_bytes_transform,
                   "fwd_ms": round(t_fwd / args.steps, 4), "inv_ms": round(t_inv / args.steps, 4),
                   "roundtrip_rel_l2": err, "d2d_copy_GBps": round(copy_gbps, 1),
                   "wall_s": round(wall, 3)},
        "passes": passes,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    del X, Y
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(n)
        except Exception as e:  # reported, never fatal
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
