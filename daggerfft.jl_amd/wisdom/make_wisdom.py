"""Regenerates the schedule cache shipped beside the library
(wisdom/<gfx arch>.txt, loaded at the first plan creation): runs the
autotuner on the device over the transforms this library is measured on and
saves the fastest schedule of every axis pass.

  python daggerfft.jl_amd/wisdom/make_wisdom.py [out.txt] [--quick]

Shapes: 512^3 ComplexF64 (full search), 128^3 ComplexF64, 256^3 R2C and C2R Float64,
384x240x210 ComplexF64, the DCT-II/III configurations, the prime-length
(Rader and Bluestein) shapes of the bench, and every distinct
request shape of the mixed serving trace (cubic 64..512, C2C and R2C, both
precisions); then the launch schedule of the 64-transform mixed batch
(grouped against lanes, with the lanes' passes tuned).  Progress is printed
per shape.
"""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    quick = "--quick" in sys.argv
    out = args[0] if args else None
    import torch
    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft import serve as S
    torch.cuda.init()
    D.tune_clear()
    jobs = [((512, 512, 512), L.C2C, L.F64, None, 2), ((128, 128, 128), L.C2C, L.F64, None, 1),
            ((256, 256, 256), L.R2C, L.F64, None, 1), ((256, 256, 256), L.C2R, L.F64, None, 1),
            ((384, 240, 210), L.C2C, L.F64, None, 1),
            ((256, 256), L.R2R, L.F64, (D.REDFT10, D.REDFT10), 1),
            ((128, 128, 128), L.R2R, L.F64, (D.REDFT10,) * 3, 1),
            # convolution passes (Rader, Bluestein): tile width and swizzle
            ((257, 257, 257), L.C2C, L.F64, None, 1), ((251, 251, 251), L.C2C, L.F64, None, 1),
            ((257, 257, 257), L.C2C, L.F32, None, 1), ((131, 131, 131), L.C2C, L.F64, None, 1),
            ((127, 127, 127), L.C2C, L.F64, None, 1), ((1009, 128, 128), L.C2C, L.F64, None, 1)]
    seen = set()
    for q in S.serving_trace():
        key = (q.shape, q.kind, q.precision)
        if key not in seen:
            seen.add(key)
            jobs.append((q.shape, q.kind, q.precision, None, 1))
    if quick:
        jobs = jobs[:8]
    t0 = time.time()
    for i, (shape, kind, prec, r2r, level) in enumerate(jobs):
        D.autotune(shape, kind=kind, sign=+1 if kind == L.C2R else -1, precision=prec, r2r_kinds=r2r, level=level)
        print("[%d/%d] %s %s %s  %.1fs  entries %d" % (i + 1, len(jobs), "x".join(map(str, shape)),
                                                     ("C2C", "R2C", "C2R", "R2R")[kind], "f64" if prec else "f32",
                                                     time.time() - t0, D.tune_entries()), flush=True)
    from daggerfft.batch import Batch, batch_config_sizes
    descs = [(shape, L.C2C, -1, prec) for shape, prec in batch_config_sizes(64, 7)]
    b = Batch(descs, flags=L.FLAG_MEASURE)
    print("batch of 64: %s schedule  %.1fs  entries %d" % ("lanes" if b.launches == 3 * len(descs) else "grouped",
                                                          time.time() - t0, D.tune_entries()), flush=True)
    b.destroy()
    if out is None:
        import subprocess
        arch = subprocess.run(["rocminfo"], capture_output=True, text=True).stdout
        arch = next((w for w in arch.split() if w.startswith("gfx")), "gfx950")
        out = os.path.join(HERE, arch + ".txt")
    D.tune_save(out)
    print("saved", D.tune_entries(), "entries to", out, flush=True)


if __name__ == "__main__":
    main()
