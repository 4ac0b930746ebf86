"""Tiles in contiguous ranges per XCD (AxisPass::swz = SWZ_XCD, -8 in a
schedule-cache line): workgroup b runs tile (b mod 8) q + min(b mod 8, r) +
b / 8 of the G = 8 q + r tiles, a bijection on [0, G).  Forced through the
schedule cache on every pass of a plan (the plan's own cache entries,
exported, with the swizzle field replaced), each result equals the same plan
with its default tile order bit for bit -- tiles are independent, only their
order changes -- and the oracle within tolerance.  Covers the catalog C2C
kernels in both layouts and precisions, Rader and Bluestein passes (tile
counts not divisible by 8), R2C and DCT-II."""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {1: 1e-12, 0: 1e-5}


def _xcd_cache(plan, tmp_path, name):
    """The plan's cache entries with every swizzle field set to -8."""
    src = os.path.join(str(tmp_path), name + ".txt")
    plan.export_wisdom(src)
    out = []
    for line in open(src):
        if line.startswith("#") or not line.strip():
            out.append(line)
            continue
        f = line.split()
        # key radices tpl tc pad variant swz
        while len(f) < 7:
            f.append("0" if len(f) < 6 else "1")
        f[6] = "-8"
        out.append(" ".join(f) + "\n")
    dst = os.path.join(str(tmp_path), name + "_xcd.txt")
    with open(dst, "w") as fh:
        fh.writelines(out)
    return dst


CASES = [
    # shape, kind, precision, r2r kinds
    ((257, 7, 5), "c2c", 1, None),     # Rader, contiguous
    ((9, 257, 5), "c2c", 1, None),     # Rader, stride n0
    ((6, 5, 251), "c2c", 0, None),     # Rader, stride n0 n1, ComplexF32
    ((131, 6, 5), "c2c", 1, None),     # Bluestein (130 is not 7-smooth)
    ((96, 80, 48), "c2c", 1, None),    # catalog kernels, all three layouts
    ((96, 80, 48), "c2c", 0, None),
    ((64, 30, 22), "r2c", 1, None),    # fused R2C pass; 22 = 2 x 11 takes a convolution pass
    ((48, 40, 36), "dct", 1, (5, 5, 5)),  # REDFT10 on every axis
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_xcd_tile_order(gpu, tmp_path, case):
    import daggerfft as D
    from daggerfft import _lib as L
    shape, kind, prec, kinds = CASES[case]
    shipped = os.path.join(os.path.dirname(D.__file__), "..", "wisdom", "gfx950.txt")
    print("xcd order", shape, kind, prec, flush=True)
    try:
        D.tune_clear()
        if kind == "c2c":
            dt = np.complex128 if prec else np.complex64
            a = O.fill(40 + case, shape, dt)
            mk = lambda: D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))  # noqa: E731
            ref = O.fftn(a.astype(np.complex128))
            oshape, odt = shape, dt
        elif kind == "r2c":
            dt = np.float64 if prec else np.float32
            a = O.fill(40 + case, shape, dt)
            mk = lambda: D.Plan(shape, L.R2C, -1, prec, D.Pencil(), (1, 2, 3))  # noqa: E731
            ref = O.rfftn(a.astype(np.float64))
            oshape, odt = (shape[0] // 2 + 1,) + shape[1:], (np.complex128 if prec else np.complex64)
        else:
            dt = np.float64 if prec else np.float32
            a = O.fill(40 + case, shape, dt)
            mk = lambda: D.Plan(shape, L.R2R, -1, prec, D.Pencil(), (1, 2, 3), r2r_kinds=kinds)  # noqa: E731
            ref = O.r2rn(a.astype(np.float64), kinds)
            oshape, odt = shape, dt
        A = D.DeviceArray.from_host(a)
        p0 = mk()
        Y0 = p0.execute(A, D.DeviceArray(oshape, odt)).to_host()
        path = _xcd_cache(p0, tmp_path, "c%d" % case)
        D.tune_clear()
        D.tune_load(path)
        D.clear_plan_cache()
        p1 = mk()
        swz = [q["swz"] for q in p1.passes()]
        assert all(s == -8 for s in swz), swz
        Y1 = p1.execute(A, D.DeviceArray(oshape, odt)).to_host()
        assert np.array_equal(Y0, Y1)
        e = O.rel_l2(Y1.astype(np.complex128 if np.iscomplexobj(Y1) else np.float64), ref)
        assert e <= TOL[prec], (shape, kind, e)
    finally:
        D.tune_clear()
        D.tune_load(shipped)
        D.clear_plan_cache()
