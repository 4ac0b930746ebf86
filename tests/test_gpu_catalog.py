"""Every kernel of the compile-time catalog (dfft_fixed_*.hip) gives the
oracle's result: each {2,3,5,7}-smooth length 16..512 along the contiguous
axis, along a strided axis with outer slices and along a strided axis
without, in both precisions, plus the fused R2C and C2R kernels of every
half-length 16..256 through each of their schedules, plus each FP64 strided
length through both of its schedules (split and full LDS exchange).  ComplexF64 rel-L2 <= 1e-12; ComplexF32
against the FP64 oracle <= 2e-6 (stated, and the worst value is printed)."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _smooth(m):
    for p in (2, 3, 5, 7):
        while m % p == 0:
            m //= p
    return m == 1


LENGTHS = [m for m in range(16, 513) if _smooth(m)]
HALF = [m for m in range(16, 257) if _smooth(m)]
TOL = {1: 1e-12, 0: 2e-6}


def _chunks(xs, k):
    n = (len(xs) + k - 1) // k
    return [xs[i * n:(i + 1) * n] for i in range(k)]


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("part", range(4))
def test_catalog_c2c_lengths(gpu, prec, part):
    import daggerfft as D
    from daggerfft import _lib as L
    D.tune_clear()
    dt = np.complex128 if prec else np.complex64
    worst = 0.0
    for n in _chunks(LENGTHS, 4)[part]:
        for shape, ax in (((n, 4, 3), 0), ((3, n, 4), 1), ((4, 3, n), 2)):
            print("C2C", "f64" if prec else "f32", shape, flush=True)
            a = O.fill(n + ax, shape, dt)
            A = D.DeviceArray.from_host(a)
            p = D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
            var = [q["variant"] for q in p.passes() if q["axis"] == ax][0]
            assert var >= 1, (shape, var)  # the catalog kernel ran
            F = p.execute(A, A.similar())
            e = O.rel_l2(F.to_host().astype(np.complex128), O.fftn(a.astype(np.complex128)))
            assert e <= TOL[prec], (shape, var, e)
            worst = max(worst, e)
            pi = D.Plan(shape, L.C2C, +1, prec, D.Pencil(), (1, 2, 3))
            B = pi.execute(F, F.similar())
            e2 = O.rel_l2(B.to_host().astype(np.complex128), a.astype(np.complex128))
            assert e2 <= TOL[prec], (shape, "inverse", e2)
    print("worst rel-L2 (%s): %.3e" % ("f64" if prec else "f32", worst))


def _force(tmp_path, line):
    import daggerfft as D
    path = os.path.join(str(tmp_path), "tune.txt")
    with open(path, "w") as f:
        f.write(line + "\n")
    D.tune_clear()
    D.tune_load(path)


@pytest.mark.parametrize("prec", [1, 0])
def test_catalog_c2r_half_lengths(gpu, tmp_path, prec):
    """The fused C2R catalog kernels (H+1 bins -> 2H reals) of every
    half-length 16..256, with the split (2) and the full (3) LDS exchange
    forced through the schedule cache: irfft of the oracle's rfft
    reconstructs the input."""
    import daggerfft as D
    from daggerfft import _lib as L
    dt = np.float64 if prec else np.float32
    cdt = np.complex128 if prec else np.complex64
    worst = 0.0
    try:
        for h in HALF:
            shape = (2 * h, 3, 2)
            a = O.fill(h + 1, shape, np.float64)
            F = O.rfftn(a).astype(cdt)
            sc = int(math.floor(math.log2(6 * h)))
            for variant in (2, 3):
                print("C2R", "f64" if prec else "f32", shape, variant, flush=True)
                _force(tmp_path, "%d/3/%d/1/%d - 0 4 0 %d" % (prec, h, sc, variant))
                p = D.Plan(shape, L.C2R, +1, prec, D.Pencil(), (1, 2, 3))
                q = [x for x in p.passes() if x["op"] == "C2R"][0]
                assert q["variant"] == variant, q
                B = p.execute(D.DeviceArray.from_host(F), D.DeviceArray(shape, dt))
                e = O.rel_l2(B.to_host().astype(np.float64), a)
                assert e <= TOL[prec], (shape, variant, e)
                worst = max(worst, e)
    finally:
        D.tune_clear()
    print("worst rel-L2 (%s): %.3e" % ("f64" if prec else "f32", worst))


@pytest.mark.parametrize("prec", [1, 0])
def test_catalog_r2c_half_lengths(gpu, tmp_path, prec):
    """The fused R2C catalog kernels of every half-length 16..256: the full
    (3) LDS exchange for all, and the split one (2, post-twiddle in two
    rounds) for H >= 100, each forced through the schedule cache."""
    import daggerfft as D
    from daggerfft import _lib as L
    dt = np.float64 if prec else np.float32
    worst = 0.0
    try:
        for h in HALF:
            shape = (2 * h, 3, 2)
            a = O.fill(h, shape, dt)
            ref = O.rfftn(a.astype(np.float64))
            sc = int(math.floor(math.log2(6 * h)))
            for variant in ((2, 3) if h >= 100 else (3,)):
                print("R2C", "f64" if prec else "f32", shape, variant, flush=True)
                _force(tmp_path, "%d/1/%d/1/%d - 0 4 0 %d" % (prec, h, sc, variant))
                p = D.Plan(shape, L.R2C, -1, prec, D.Pencil(), (1, 2, 3))
                assert p.passes()[0]["variant"] == variant and p.passes()[0]["op"] == "R2C"
                F = p.execute(D.DeviceArray.from_host(a), D.DeviceArray((h + 1, 3, 2), np.complex128 if prec else np.complex64))
                e = O.rel_l2(F.to_host().astype(np.complex128), ref)
                assert e <= TOL[prec], (shape, variant, e)
                worst = max(worst, e)
    finally:
        D.tune_clear()
    print("worst rel-L2 (%s): %.3e" % ("f64" if prec else "f32", worst))


def test_catalog_f64_strided_both_schedules(gpu, tmp_path):
    """FP64 strided lines have a split-LDS (2) and a full-LDS (3) kernel;
    each is forced through the schedule cache for every length."""
    import daggerfft as D
    from daggerfft import _lib as L
    for variant in (2, 3):
        for n in LENGTHS:
            if n & (n - 1) == 0 and n >= 64 and variant == 3:
                continue  # powers of two >= 64 have no variant-3 kernel
            shape = (2, n, 3)
            sc = int(math.floor(math.log2(int(np.prod(shape)))))
            path = os.path.join(str(tmp_path), "tune.txt")
            with open(path, "w") as f:
                f.write("1/0/%d/2/%d - 0 8 0 %d\n" % (n, sc, variant))
            D.tune_clear()
            D.tune_load(path)
            try:
                p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
                got = [q["variant"] for q in p.passes() if q["axis"] == 1][0]
                assert got == variant, (n, got)
                a = O.fill(n, shape, np.complex128)
                F = p.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, np.complex128))
                e = O.rel_l2(F.to_host(), O.fftn(a))
                assert e <= 1e-12, (n, variant, e)
            finally:
                D.tune_clear()


@pytest.mark.parametrize("prec", [1, 0])
def test_catalog_pow2_alternative_radix_lists(gpu, tmp_path, prec):
    """Powers of two have two split-LDS kernels per length (radix-8 lists and
    fewest-stage lists such as 16*16 for 256); the schedule cache names the
    list.  Each is forced for both layouts and checked against the oracle."""
    import daggerfft as D
    from daggerfft import _lib as L
    lists = {64: "16,4", 128: "16,8", 256: "16,16", 512: "16,16,2", 1024: "16,16,4"}
    dt = np.complex128 if prec else np.complex64
    for n, rad in lists.items():
        for shape, ax, layout in (((n, 4, 3), 0, 1), ((3, n, 4), 1, 2)):
            sc = int(math.floor(math.log2(int(np.prod(shape)))))
            path = os.path.join(str(tmp_path), "tune.txt")
            with open(path, "w") as f:
                f.write("%d/0/%d/%d/%d %s 0 %d 0 2\n" % (prec, n, layout, sc, rad, 1 if layout == 1 else 4))
            D.tune_clear()
            D.tune_load(path)
            try:
                p = D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
                q = [x for x in p.passes() if x["axis"] == ax][0]
                assert q["variant"] == 2 and q["radices"] == rad.replace(",", "x"), q
                a = O.fill(n + ax, shape, dt)
                F = p.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, dt))
                e = O.rel_l2(F.to_host().astype(np.complex128), O.fftn(a.astype(np.complex128)))
                assert e <= TOL[prec], (n, ax, e)
            finally:
                D.tune_clear()


@pytest.mark.parametrize("variant", [2])
@pytest.mark.parametrize("prec", [1, 0])
def test_catalog_dct_pow2(gpu, prec, variant, tmp_path):
    """DCT-II / DCT-III catalog kernels through the n-point FFT
    (DFFT_FLAG_DCT_FULL_LENGTH; split 16-value schedule, forced through the
    schedule cache) of the powers of two 16..1024, along the contiguous axis
    (paired loads and stores) and a strided axis, on 2 and 6 lines, against
    the oracle."""
    import daggerfft as D
    from daggerfft import _lib as L
    dt = np.float64 if prec else np.float32
    tol = 1e-12 if prec else 1e-5
    cases = []
    for n in (16, 32, 64, 128, 256, 512, 1024):
        for shape, ax, layout in (((n, 3, 2), 0, 1), ((n, 2, 1), 0, 1), ((2, n, 3), 1, 2)):
            lines = int(np.prod(shape)) // n
            cases.append((n, shape, ax, layout, int(math.floor(math.log2(lines * n)))))
    path = str(tmp_path / "dct.txt")
    with open(path, "w") as f:
        for n, shape, ax, layout, sc in cases:
            for op in (5, 6):  # DCT2, DCT3
                f.write("%d/%d/%d/%d/%d - 0 0 0 %d 1\n" % (prec, op, n, layout, sc, variant))
    D.tune_clear()
    D.tune_load(path)
    try:
        for n, shape, ax, layout, sc in cases:
            a = O.fill(n + ax, shape, dt)
            for kind, okind in ((D.REDFT10, O.REDFT10), (D.REDFT01, O.REDFT01)):
                p = D.Plan(shape, L.R2R, -1, prec, D.Pencil(), (ax + 1,), r2r_kinds=(kind,),
                           flags=L.FLAG_DCT_FULL_LENGTH)
                q = p.passes()[0]
                assert q["variant"] == variant and q["op"] in ("DCT2", "DCT3"), q
                G = p.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, dt))
                e = O.rel_l2(G.to_host().astype(np.float64), O.r2rn(a.astype(np.float64), okind, (ax + 1,)))
                assert e <= tol, (n, shape, ax, kind, e)
    finally:
        D.tune_clear()


@pytest.mark.parametrize("prec", [1, 0])
def test_catalog_dct_half_length(gpu, prec):
    """REDFT10 / REDFT01 of n = 16..1024 reals through the n/2-point complex
    FFT (DFFT_FLAG_DCT_HALF_LENGTH): along the contiguous axis (quad loads
    and stores across mirror lanes) and the strided axes, on 2 and 6 lines,
    against the oracle; and equal, to rounding, to the n-point path.  Without
    a flag a plan takes the half-length path from n = 256 (REDFT10) and
    n = 512 (REDFT01) in arrays of 4M elements or more."""
    import daggerfft as D
    from daggerfft import _lib as L
    dt = np.float64 if prec else np.float32
    tol = 1e-12 if prec else 1e-5
    for n in (16, 32, 64, 128, 256, 512, 1024):
        for shape, ax in (((n, 3, 2), 0), ((n, 2, 1), 0), ((2, n, 3), 1), ((3, 2, n), 2)):
            a = O.fill(3 * n + ax, shape, dt)
            for kind, okind, op in ((D.REDFT10, O.REDFT10, "DCT2H"), (D.REDFT01, O.REDFT01, "DCT3H")):
                p = D.Plan(shape, L.R2R, -1, prec, D.Pencil(), (ax + 1,), r2r_kinds=(kind,),
                           flags=L.FLAG_DCT_HALF_LENGTH)
                q = p.passes()[0]
                assert q["op"] == op and q["L"] == n // 2 and q["nlog"] == n, q
                d = D.Plan(shape, L.R2R, -1, prec, D.Pencil(), (ax + 1,), r2r_kinds=(kind,)).passes()[0]
                assert d["op"] == op[:4], d  # small arrays keep the n-point FFT
                G = p.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, dt)).to_host()
                e = O.rel_l2(G.astype(np.float64), O.r2rn(a.astype(np.float64), okind, (ax + 1,)))
                assert e <= tol, (n, shape, ax, op, e)
                f = D.Plan(shape, L.R2R, -1, prec, D.Pencil(), (ax + 1,), r2r_kinds=(kind,),
                           flags=L.FLAG_DCT_FULL_LENGTH)
                F = f.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, dt)).to_host()
                assert O.rel_l2(G.astype(np.float64), F.astype(np.float64)) <= tol
    for shape, ax, ok2, ok3 in (((256, 256, 64), 0, True, False), ((512, 128, 64), 0, True, True),
                                ((128, 256, 128), 0, False, False), ((256, 256, 32), 0, False, False)):
        for kind, op, ok in ((D.REDFT10, "DCT2", ok2), (D.REDFT01, "DCT3", ok3)):
            d = D.Plan(shape, L.R2R, -1, prec, D.Pencil(), (ax + 1,), r2r_kinds=(kind,)).passes()[0]
            assert d["op"] == (op + "H" if ok else op), (shape, d)


@pytest.mark.parametrize("swz", [97, 1021])
def test_swizzled_tile_order(gpu, tmp_path, swz):
    """A strided pass whose workgroups take the tiles in swizzled order
    (tile = block * swz mod tiles, forced through the schedule cache) gives
    the same result as the in-order pass and the oracle."""
    import daggerfft as D
    from daggerfft import _lib as L
    shape = (64, 48, 128)
    a = O.fill(31, shape, np.complex128)
    sc = int(math.floor(math.log2(int(np.prod(shape)))))
    path = os.path.join(str(tmp_path), "tune.txt")
    with open(path, "w") as f:
        f.write("1/0/128/0/%d - 0 8 0 2 %d\n" % (sc, swz))   # axis 3: plane layout
        f.write("1/0/48/2/%d - 0 8 0 2 %d\n" % (sc, swz))    # axis 2: outer slices
    D.tune_clear()
    D.tune_load(path)
    try:
        p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
        sw = {q["axis"]: q["swz"] for q in p.passes()}
        assert sw[2] == swz and sw[1] == swz and sw[0] == 1, sw
        F = p.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, np.complex128))
    finally:
        D.tune_clear()
    ref = D.fft(D.DeviceArray.from_host(a)).to_host()
    assert np.array_equal(F.to_host(), ref)
    assert O.rel_l2(F.to_host(), O.fftn(a)) <= 1e-12
