"""Benchmark: 512^3 ComplexF64 3D C2C, forward + inverse, one GPU.

Methodology of the reference benchmark (benchmarks/fftmpigpu.jl:76-105):
warm-up, then timed iterations of fft! followed by ifft!;
time_per_transform = (t_fft + t_ifft) / 2 and GFLOP/s = 5 N log2(N) 1e-9 / t.

One "step" is one forward plus one inverse transform of the resident
512x512x512 ComplexF64 block (inputs generated on device by the SplitMix64
counter generator, seed 3).  Both decompositions are measured:
  * Pencil: one launch per axis pass; every pass is bracketed by HIP events
    on the stream it runs on (per-pass GB/s against the 8 TB/s HBM peak);
  * Slab: the 2D stage software-pipelined over chunks of z-planes in grouped
    launches, then the third axis; each transform bracketed by HIP events.
`value` is the faster decomposition (named in config.decomp).

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
    ap.add_argument("--decomp", choices=["pencil", "slab", "both"], default="both")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--profile-steps", action="store_true", help="no per-pass events (for rocprof runs)")
    return ap.parse_args()


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs, calibrated on kernels of known size)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_512cubed_c2c_f64.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch "):
                name, val = line[len("traffic_bytes_per_launch "):].rsplit("=", 1)
                if name.strip() == kernel.replace(" ", ""):
                    return float(val)
    except OSError:
        pass
    return None


def cpu_baseline(n):
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
            "(per-axis batched FFTs in the reference's stage order), time per transform %.3f s" % (n, t)}


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
    A = D.rand(shape, np.complex128, seed=3)
    X = A.similar()
    Y = A.similar()
    Ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    results = {}

    def roundtrip_err():
        return float(torch.linalg.vector_norm(Y.tensor - A.tensor) / torch.linalg.vector_norm(A.tensor))

    # ---------------- Pencil: one launch per axis pass ----------------
    if args.decomp in ("pencil", "both"):
        pf = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
        pi = D.Plan(shape, L.C2C, +1, L.F64, D.Pencil(), (1, 2, 3))
        fpasses, ipasses = pf.passes(), pi.passes()
        keys = [("f", i) for i in range(len(fpasses))] + [("i", i) for i in range(len(ipasses))]

        def step(ev=None):
            for i in range(len(fpasses)):
                if ev:
                    ev[("f", i)][0].record()
                pf.execute_pass(i, A, X)
                if ev:
                    ev[("f", i)][1].record()
            for i in range(len(ipasses)):
                if ev:
                    ev[("i", i)][0].record()
                pi.execute_pass(i, X, Y)
                if ev:
                    ev[("i", i)][1].record()

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        s0, s1 = Ev(), Ev()
        evs = []
        s0.record()
        for _ in range(args.steps):
            ev = None if args.profile_steps else {k: (Ev(), Ev()) for k in keys}
            step(ev)
            evs.append(ev)
        s1.record()
        torch.cuda.synchronize()
        ms_step = s0.elapsed_time(s1) / args.steps
        passes = []
        for k in keys:
            desc = (fpasses if k[0] == "f" else ipasses)[k[1]]
            avg = None if args.profile_steps else sum(e[k][0].elapsed_time(e[k][1]) for e in evs) / args.steps
            gbps = desc["bytes"] / (avg / 1e3) / 1e9 if avg else None
            passes.append({"dir": "fwd" if k[0] == "f" else "inv", "axis": desc["axis"] + 1, "op": desc["op"],
                           "L": desc["L"], "radices": desc["radices"], "TPL": desc["TPL"], "TC": desc["TC"],
                           "variant": desc["variant"], "kernel": desc["kernel"],
                           "bytes": desc["bytes"], "avg_ms": round(avg, 4) if avg else None,
                           "GBps": round(gbps, 1) if gbps else None,
                           "frac": round(gbps / HBM_PEAK_GBPS, 4) if gbps else None})
        results["pencil"] = {"ms_step": ms_step, "passes": passes, "err": roundtrip_err(),
                             "bytes": pf.algorithmic_bytes()}

    # ---------------- Slab: software-pipelined grouped launches ----------------
    if args.decomp in ("slab", "both"):
        sf = D.Plan(shape, L.C2C, -1, L.F64, D.Slab(), (1, 2, 3))
        si = D.Plan(shape, L.C2C, +1, L.F64, D.Slab(), (1, 2, 3))
        sched, launches = sf.schedule()
        for _ in range(args.warmup):
            sf.execute(A, X)
            si.execute(X, Y)
        torch.cuda.synchronize()
        s0, s1 = Ev(), Ev()
        evs = []
        s0.record()
        for _ in range(args.steps):
            e = (Ev(), Ev(), Ev())
            e[0].record()
            sf.execute(A, X)
            e[1].record()
            si.execute(X, Y)
            e[2].record()
            evs.append(e)
        s1.record()
        torch.cuda.synchronize()
        ms_step = s0.elapsed_time(s1) / args.steps
        fwd = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps
        inv = sum(e[1].elapsed_time(e[2]) for e in evs) / args.steps
        results["slab"] = {"ms_step": ms_step, "fwd_ms": fwd, "inv_ms": inv, "err": roundtrip_err(),
                           "schedule": sched, "launches": launches, "bytes": sf.algorithmic_bytes()}

    # measured device-to-device copy ceiling on the same buffers
    torch.cuda.synchronize()
    c0, c1 = Ev(), Ev()
    X.tensor.copy_(A.tensor)
    c0.record()
    for _ in range(5):
        X.tensor.copy_(A.tensor)
    c1.record()
    torch.cuda.synchronize()
    copy_gbps = 2 * A.nbytes * 5 / (c0.elapsed_time(c1) / 1e3) / 1e9

    best = min(results, key=lambda k: results[k]["ms_step"])
    r = results[best]
    ms_step = r["ms_step"]
    t_transform_s = ms_step / 2 / 1e3
    gflops = 5 * N * math.log2(N) * 1e-9 / t_transform_s

    # Dominant kernel: the slowest axis pass of the per-pass schedule (Pencil and
    # Slab run the same axis-pass kernels; Slab only groups them differently).
    roofline = None
    if "pencil" in results and not args.profile_steps:
        dom = max(results["pencil"]["passes"], key=lambda p: p["avg_ms"])
        roofline = {"bound": "hbm", "achieved": dom["GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": dom["frac"], "traffic": pmc_traffic(dom["kernel"]),
                    "kernel": "%s (axis %d, %s)" % (dom["kernel"], dom["axis"], dom["dir"]),
                    "bytes_per_launch": dom["bytes"], "avg_ms": dom["avg_ms"]}

    cfg = {"workload": "3D C2C ComplexF64 %dx%dx%d forward+inverse" % (n, n, n), "decomp": best,
           "dims": [1, 2, 3], "step": "fft then ifft", "time_per_transform_ms": round(ms_step / 2, 4),
           "hbm_frac_per_transform": round(r["bytes"] / t_transform_s / 1e9 / HBM_PEAK_GBPS, 4),
           "algorithmic_bytes_per_transform": r["bytes"], "roundtrip_rel_l2": r["err"],
           "d2d_copy_GBps": round(copy_gbps, 1)}
    for k, v in results.items():
        cfg["%s_ms_per_transform" % k] = round(v["ms_step"] / 2, 4)
        cfg["%s_GFLOPs" % k] = round(5 * N * math.log2(N) * 1e-9 / (v["ms_step"] / 2 / 1e3), 1)
    if "slab" in results:
        cfg["slab_schedule"] = results["slab"]["schedule"]
        cfg["slab_launches_per_transform"] = results["slab"]["launches"]
        cfg["slab_roundtrip_rel_l2"] = results["slab"]["err"]
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
        "config": cfg,
        "passes": results.get("pencil", {}).get("passes"),
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
