"""Long lines: the power-of-two lengths 2048..16384 (dfft_fixed_long.hip) and
every {2,3,5,7}-smooth length 1025..4096 (dfft_fixed_mid*.hip) run
compile-time catalog kernels (16 values per thread, split-LDS exchange) along
the contiguous axis, a strided axis with outer slices and a strided axis
without, in both precisions; FP64 4096 through both of its radix lists.  The
fused C2R (split exchange) of every smooth half-length 257..1024
(dfft_fixed_real*.hip) likewise, with the R2C of the same lines (runtime
engine).  Full-size long-line shapes whose tuned schedules the shipped cache
holds, every element against the oracle.
ComplexF64 rel-L2 <= 1e-12; ComplexF32 against the FP64 oracle <= 2e-6."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = {1: 1e-12, 0: 2e-6}
LONG = [2048, 4096, 8192, 16384]


def _smooth(m):
    for p in (2, 3, 5, 7):
        while m % p == 0:
            m //= p
    return m == 1


MID = [m for m in range(1025, 4097) if _smooth(m) and m & (m - 1)]
HALF_LONG = [h for h in range(257, 1025) if _smooth(h)]


def _case(D, L, shape, prec, ax):
    dt = np.complex128 if prec else np.complex64
    a = O.fill(shape[ax] + ax, shape, dt)
    p = D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
    q = [x for x in p.passes() if x["axis"] == ax][0]
    F = p.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, dt))
    e = O.rel_l2(F.to_host().astype(np.complex128), O.fftn(a.astype(np.complex128)))
    pi = D.Plan(shape, L.C2C, +1, prec, D.Pencil(), (1, 2, 3))
    B = pi.execute(F, F.similar())
    e2 = O.rel_l2(B.to_host().astype(np.complex128), a.astype(np.complex128))
    return q, e, e2


@pytest.mark.parametrize("n", LONG)
@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
def test_long_lines_catalog(gpu, n, prec):
    import daggerfft as D
    from daggerfft import _lib as L
    D.tune_clear()
    for shape, ax in (((n, 3, 2), 0), ((3, n, 2), 1), ((2, 3, n), 2)):
        print("long", n, "f64" if prec else "f32", shape, flush=True)
        q, e, e2 = _case(D, L, shape, prec, ax)
        assert q["variant"] == 2 and q["L"] == n, q
        assert e <= TOL[prec] and e2 <= TOL[prec], (shape, e, e2)
    if prec == 1 and n == 4096:
        assert q["radices"] == "8x8x8x8", q  # the default list of FP64 4096 (16x16x16 spills)


@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
@pytest.mark.parametrize("part", range(4))
def test_mid_lengths_catalog(gpu, prec, part):
    import daggerfft as D
    from daggerfft import _lib as L
    D.tune_clear()
    worst = 0.0
    for n in MID[part::4]:
        for shape, ax in (((n, 3, 2), 0), ((3, n, 2), 1), ((2, 3, n), 2)):
            print("mid", n, "f64" if prec else "f32", shape, flush=True)
            q, e, e2 = _case(D, L, shape, prec, ax)
            assert q["variant"] == 2 and q["L"] == n, q
            assert e <= TOL[prec] and e2 <= TOL[prec], (shape, e, e2)
            worst = max(worst, e, e2)
    print("worst rel-L2 (%s): %.3e" % ("f64" if prec else "f32", worst))


@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
def test_long_half_lengths_r2c_c2r(gpu, prec):
    import daggerfft as D
    from daggerfft import _lib as L
    D.tune_clear()
    rdt = np.float64 if prec else np.float32
    cdt = np.complex128 if prec else np.complex64
    worst = 0.0
    for h in HALF_LONG:
        shape = (2 * h, 3, 2)
        print("R2C/C2R", h, "f64" if prec else "f32", flush=True)
        a = O.fill(h, shape, rdt)
        ref = O.rfftn(a.astype(np.float64))
        pr = D.Plan(shape, L.R2C, -1, prec, D.Pencil(), (1, 2, 3))
        q = pr.passes()[0]
        assert q["op"] == "R2C", q
        F = pr.execute(D.DeviceArray.from_host(a), D.DeviceArray((h + 1, 3, 2), cdt))
        e = O.rel_l2(F.to_host().astype(np.complex128), ref)
        pc = D.Plan(shape, L.C2R, +1, prec, D.Pencil(), (1, 2, 3))
        qc = [x for x in pc.passes() if x["op"] == "C2R"][0]
        assert qc["variant"] == 2, qc
        B = pc.execute(D.DeviceArray.from_host(ref.astype(cdt)), D.DeviceArray(shape, rdt))
        e2 = O.rel_l2(B.to_host().astype(np.float64), a.astype(np.float64))
        assert e <= TOL[prec] and e2 <= TOL[prec], (h, e, e2)
        worst = max(worst, e, e2)
    print("worst rel-L2 (%s): %.3e" % ("f64" if prec else "f32", worst))


def test_f64_4096_both_radix_lists(gpu, tmp_path):
    """FP64 4096 holds 8x8x8x8 (default) and 16x16x16 (forced through the
    schedule cache, key prec/op/L/layout/size class)."""
    import daggerfft as D
    from daggerfft import _lib as L
    path = os.path.join(str(tmp_path), "tune.txt")
    try:
        for shape, ax, layout in (((4096, 3, 2), 0, 1), ((3, 4096, 2), 1, 2), ((2, 3, 4096), 2, 0)):
            sc = int(math.floor(math.log2(6 * 4096)))
            with open(path, "w") as f:
                f.write("1/0/4096/%d/%d 16,16,16 0 0 0 2 1\n" % (layout, sc))
            D.tune_clear()
            D.tune_load(path)
            q, e, e2 = _case(D, L, shape, 1, ax)
            assert q["radices"] == "16x16x16" and q["variant"] == 2, q
            assert e <= 1e-12 and e2 <= 1e-12, (shape, e, e2)
    finally:
        D.tune_clear()


SHIPPED_LONG = [
    # shape, axis whose line is long; the shipped schedule cache holds tuned
    # entries for these sizes (a strided 4096-point pass runs one contiguous
    # range of tiles per XCD)
    ((256, 4096, 128), 1),
    ((128, 128, 4096), 2),
    ((16384, 64, 64), 0),
]


@pytest.mark.parametrize("case", range(len(SHIPPED_LONG)))
def test_long_lines_shipped_schedules_full_array(gpu, case):
    """Full-size long-line transforms with the shipped schedule cache
    (64M-134M points): every output element of the forward transform against
    the oracle, and the inverse back to the input."""
    import daggerfft as D
    from daggerfft import _lib as L
    shape, ax = SHIPPED_LONG[case]
    print("long, shipped schedules", shape, flush=True)
    D.tune_clear()
    D.tune_load(os.path.join(os.path.dirname(D.__file__), "..", "wisdom", "gfx950.txt"))
    a = O.fill(7 + case, shape, np.complex128)
    p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    q = [x for x in p.passes() if x["axis"] == ax][0]
    assert q["L"] == shape[ax], q
    if ax > 0:
        assert q["swz"] == -8, q
    A = D.DeviceArray.from_host(a)
    F = p.execute(A, D.DeviceArray(shape, np.complex128))
    del A
    y = F.to_host()
    e = O.rel_l2(y, O.fftn(a))
    del y
    pi = D.Plan(shape, L.C2C, +1, L.F64, D.Pencil(), (1, 2, 3))
    B = pi.execute(F, F.similar())
    del F
    e2 = O.rel_l2(B.to_host(), a)
    print("rel-L2 fwd %.3e, fwd+inv %.3e" % (e, e2))
    assert e <= 1e-12 and e2 <= 1e-12, (shape, e, e2)


def test_longest_lines(gpu):
    """16384 points is the longest line (one LDS tile, 1024 threads); a longer
    power of two is refused with DFFT_ERR_UNSUPPORTED_SIZE."""
    import daggerfft as D
    from daggerfft import _lib as L
    for prec in (1, 0):
        D.Plan((16384, 1, 1), L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
        with pytest.raises(D.DFFTError) as ei:
            D.Plan((32768, 1, 1), L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
        assert ei.value.status == L.DFFT_ERR_UNSUPPORTED_SIZE
