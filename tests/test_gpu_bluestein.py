"""Bluestein (chirp-z) passes: line lengths with a prime factor above 7,
which FFTW and cuFFT accept and the Stockham catalog cannot factor.  Each
such axis runs as two M-point FFTs of a cyclic convolution in LDS (M >= 2n-1),
around the same fused load/store maps as every other pass: C2C passes, and
R2C/C2R along the contiguous axis, on the compile-time engine (M one of 2^k,
3*2^k, 5*2^k); the others -- and any of these when the schedule cache names
it -- on the runtime engine (M 7-smooth).

Parity with the oracle (whose recursive FFT ends in direct DFTs of the prime
factors, pinned against scipy in test_oracle.py): ComplexF64/Float64 rel-L2
<= 1e-12; ComplexF32/Float32 against the FP64 oracle <= 2e-5."""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL64 = 1e-12
TOL32 = 2e-5


def _dev(a):
    import daggerfft as D
    return D.DeviceArray.from_host(a)


C2C_SHAPES = [(11, 13, 17), (97, 6, 5), (8, 101, 4), (6, 5, 127), (1009, 4, 3), (4, 3, 1031), (22, 26, 34),
              (2003, 2, 2), (2, 4093, 1), (4093, 2, 1)]


@pytest.mark.parametrize("shape", C2C_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
def test_c2c_prime_axes_vs_oracle(gpu, shape, prec):
    import daggerfft as D
    dt = np.complex128 if prec else np.complex64
    tol = TOL64 if prec else TOL32
    a = O.fill(21, shape, dt)
    ref_in = a.astype(np.complex128)
    F = D.fft(_dev(a), decomp=D.Pencil(), dims=(1, 2, 3))
    e_f = O.rel_l2(F.to_host().astype(np.complex128), O.fftn(ref_in))
    B = D.ifft(_dev(a), decomp=D.Pencil(), dims=(1, 2, 3))
    e_b = O.rel_l2(B.to_host().astype(np.complex128), O.fftn(ref_in, inverse=True))
    R = D.ifft(F, decomp=D.Pencil(), dims=(1, 2, 3))
    e_r = O.rel_l2(R.to_host().astype(np.complex128), ref_in)
    assert e_f <= tol and e_b <= tol and e_r <= tol, (e_f, e_b, e_r)


@pytest.mark.parametrize("shape", [(11, 13, 17), (30, 97, 8)], ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("dims", [(1, 2, 3), (3, 1, 2), (2,), (1, 3)], ids=lambda d: "".join(map(str, d)))
def test_c2c_prime_dims_and_slab(gpu, shape, dims):
    import daggerfft as D
    a = O.fill(22, shape, np.complex128)
    ref = O.fftn(a, dims)
    for dec in (D.Pencil(), D.Slab()):
        if isinstance(dec, D.Slab) and len(dims) != 3:
            continue
        assert O.rel_l2(D.fft(_dev(a), decomp=dec, dims=dims).to_host(), ref) <= TOL64


@pytest.mark.parametrize("shape", [(22, 13, 6), (13, 11, 4), (2018, 3, 2), (26, 34, 38), (3, 2, 1)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
def test_rfft_irfft_prime_lengths(gpu, shape, prec):
    """Even n1 with a prime half (packed R2C/C2R around the Bluestein core),
    odd prime n1 (full-length R2C/C2R), prime edges on the other axes."""
    import daggerfft as D
    rdt = np.float64 if prec else np.float32
    tol = TOL64 if prec else TOL32
    x = O.fill(23, shape, rdt)
    F = D.rfft(_dev(x), decomp=D.Pencil(), dims=(1, 2, 3))
    ref = O.rfftn(x.astype(np.float64))
    assert O.rel_l2(F.to_host().astype(np.complex128), ref) <= tol
    y = D.irfft(_dev(ref.astype(np.complex128 if prec else np.complex64)), shape[0], decomp=D.Pencil(),
                dims=(1, 2, 3))
    assert O.rel_l2(y.to_host().astype(np.float64), x.astype(np.float64)) <= tol


KINDS = ["DHT", "REDFT00", "REDFT01", "REDFT10", "REDFT11", "RODFT00", "RODFT01", "RODFT10", "RODFT11"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("shape", [(11, 13), (12, 10, 3), (97, 5)], ids=lambda s: "x".join(map(str, s)))
def test_r2r_kinds_bluestein(gpu, kind, shape):
    """Every r2r kind where its FFT length has a prime factor above 7
    (REDFT00 of n = 12 is a 22-point FFT, RODFT00 of n = 10 too)."""
    import daggerfft as D
    a = O.fill(24, shape, np.float64)
    dims = tuple(range(1, len(shape) + 1))
    kinds = tuple(getattr(D, kind) for _ in dims)
    G = D.fft(_dev(a), decomp=D.Pencil(), dims=dims, transforms=D.R2R(kinds))
    ref = O.r2rn(a, tuple(getattr(O, kind) for _ in dims), dims)
    assert O.rel_l2(G.to_host(), ref) <= TOL64


# 263 and 773 are primes whose n-1 is not 7-smooth (so not Rader passes)
MIXED_M_EDGES = [263, 513, 773, 1025, 1537, 2049]


@pytest.mark.parametrize("n", MIXED_M_EDGES)
@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
def test_c2c_compile_time_convolution_lengths(gpu, n, prec):
    """Edges whose cheapest compile-time convolution length is a power of two
    or 3*2^k / 5*2^k (263 -> 640, 513 -> 1280, 1025 -> 2560, ...), along
    each layout, against the oracle."""
    import daggerfft as D
    from daggerfft import _lib as L
    dt = np.complex128 if prec else np.complex64
    tol = TOL64 if prec else TOL32
    for shape, ax in (((n, 3, 2), 0), ((3, n, 2), 1), ((2, 3, n), 2)):
        a = O.fill(n + ax, shape, dt)
        p = D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
        q = [x for x in p.passes() if x["axis"] == ax][0]
        assert q["kernel"].startswith("k_blue_fixed") and q["bluestein"] >= 2 * n - 1, q
        F = p.execute(_dev(a), D.DeviceArray(shape, dt))
        e = O.rel_l2(F.to_host().astype(np.complex128), O.fftn(a.astype(np.complex128)))
        assert e <= tol, (shape, q["bluestein"], e)


def test_plan_describes_bluestein_passes(gpu):
    """C2C passes, and real ops along the contiguous axis, run the
    compile-time engine over the cheapest catalog convolution length; the
    packed R2C op runs its full LDS exchange; a real op along a strided axis
    runs the runtime engine over a 7-smooth length."""
    import daggerfft as D
    from daggerfft import _lib as L
    p = D.Plan((23, 16, 1013), L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    by_axis = {q["axis"]: q for q in p.passes()}
    assert by_axis[0]["bluestein"] == 48 and by_axis[0]["kernel"].startswith("k_blue_fixed<double,0,true,true")
    assert D.Plan((263, 2, 2), L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3)).passes()[0]["bluestein"] == 640
    assert by_axis[1]["bluestein"] == 0
    assert by_axis[2]["bluestein"] == 2048 and by_axis[2]["kernel"].startswith("k_blue_fixed<double,0,false,true")
    # primes whose n-1 is 7-smooth run Rader passes over n-1 where that FFT
    # has no more stages than the Bluestein one: 11 (10, one stage against
    # 16x2) and 257 (8x8x4 against 16x8x5) do; 1009 in FP64 (1008 needs four
    # stages of radices up to 10 against 16x16x8) does not
    p = D.Plan((11, 16, 1009), L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    by_axis = {q["axis"]: q for q in p.passes()}
    assert by_axis[0]["rader"] == 10 and by_axis[0]["kernel"].startswith("k_rader_fixed<double,true")
    assert by_axis[2]["bluestein"] == 2048 and by_axis[2]["kernel"].startswith("k_blue_fixed<double,0,false")
    assert D.Plan((257, 2, 2), L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3)).passes()[0]["rader"] == 256
    r = D.Plan((22, 16, 8), L.R2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    q = r.passes()[0]
    assert q["op"] == "R2C" and q["bluestein"] == 32 and q["kernel"].startswith("k_blue_fixed<double,1,true,false")
    c = D.Plan((13, 16, 8), L.C2R, +1, L.F32, D.Pencil(), (1, 2, 3))
    q = [x for x in c.passes() if x["op"] == "C2R_FULL"][0]
    assert q["kernel"].startswith("k_blue_fixed<float,4,true,true")
    s = D.Plan((16, 22, 8), L.R2C, -1, L.F64, D.Pencil(), (2, 1, 3))
    q = s.passes()[0]
    assert q["op"] == "R2C" and q["axis"] == 1 and q["kernel"].startswith("k_axis_blue<double,1,false")


@pytest.mark.parametrize("shape", [(11, 13, 17), (97, 6, 5), (4, 3, 1031), (2003, 2, 2)], ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
def test_c2c_runtime_bluestein_forced(gpu, tmp_path, shape, prec):
    """The runtime-engine Bluestein C2C pass, named through the schedule
    cache (variant 0) for every axis, against the oracle."""
    import math
    import os
    import daggerfft as D
    from daggerfft import _lib as L
    lines = []
    for ax in range(3):
        n = shape[ax]
        layout = 1 if ax == 0 else (2 if ax == 1 and shape[2] > 1 else 0)
        sc = int(math.floor(math.log2(max(1, np.prod(shape)))))
        lines.append("%d/0/%d/%d/%d - 0 0 0 0 1" % (prec, n, layout, sc))
    path = os.path.join(str(tmp_path), "tune.txt")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    dt = np.complex128 if prec else np.complex64
    tol = TOL64 if prec else TOL32
    try:
        D.tune_clear()
        D.tune_load(path)
        p = D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
        kinds = [q["kernel"].split("<")[0] for q in p.passes() if q["bluestein"]]
        assert kinds and all(k == "k_axis_blue" for k in kinds), p.passes()
        a = O.fill(26, shape, dt)
        F = p.execute(_dev(a), D.DeviceArray(shape, dt))
        assert O.rel_l2(F.to_host().astype(np.complex128), O.fftn(a.astype(np.complex128))) <= tol
    finally:
        D.tune_clear()


def test_plane_pipeline_and_graph_with_bluestein(gpu):
    """A Slab plane-pipeline request falls back to one launch per pass (the
    grouped kernel has no Bluestein tile); graph replay runs the same kernels."""
    import daggerfft as D
    from daggerfft import _lib as L
    shape = (26, 13, 11)
    a = O.fill(25, shape, np.complex128)
    ref = O.fftn(a)
    p = D.Plan(shape, L.C2C, -1, L.F64, D.Slab(), (1, 2, 3), flags=L.FLAG_PLANE_PIPELINE)
    assert p.schedule()[0] == "passes"
    out = D.DeviceArray(shape, np.complex128)
    p.execute(_dev(a), out)
    assert O.rel_l2(out.to_host(), ref) <= TOL64
    g = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3), flags=L.FLAG_GRAPH)
    X = _dev(a)
    for _ in range(3):
        out2 = D.DeviceArray(shape, np.complex128)
        g.execute(X, out2)
        assert O.rel_l2(out2.to_host(), ref) <= TOL64


def test_batch_with_prime_lengths(gpu):
    """The default batch schedule runs transforms with a Bluestein axis on
    per-transform plans (lanes); the grouped schedule refuses them."""
    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft.batch import Batch
    descs = [((11, 16, 8), L.C2C, -1, 1), ((32, 13, 17), L.C2C, -1, 0), ((22, 10, 9), L.R2C, -1, 1)]
    b = Batch(descs)
    ins, outs, refs = [], [], []
    for t, (shape, kind, sign, prec) in enumerate(descs):
        if kind == L.C2C:
            a = O.fill(30 + t, shape, np.complex128 if prec else np.complex64)
            outs.append(D.DeviceArray(shape, a.dtype))
            refs.append(O.fftn(a.astype(np.complex128)))
        else:
            a = O.fill(30 + t, shape, np.float64)
            outs.append(D.DeviceArray((shape[0] // 2 + 1,) + shape[1:], np.complex128))
            refs.append(O.rfftn(a))
        ins.append(_dev(a))
    b.execute(ins, outs)
    for (shape, kind, sign, prec), Y, ref in zip(descs, outs, refs):
        assert O.rel_l2(Y.to_host().astype(np.complex128), ref) <= (TOL64 if prec else TOL32)
    with pytest.raises(D.DFFTError) as ei:
        Batch(descs[:2], flags=L.FLAG_BATCH_GROUPED)
    assert ei.value.status == L.DFFT_ERR_UNSUPPORTED_SIZE


def test_lengths_beyond_one_tile_are_refused(gpu):
    """Bluestein needs M >= 2n-1 points of one line in LDS with at most 1024
    threads.  C2C runs up to n = 8192 (the compile-time engine's M = 16384):
    8191 (prime) runs and matches the oracle, 8209 (prime) is refused.  The
    runtime engine (the other ops) takes M <= 8192: R2C of 2 x 4099 points
    (a 4099-point half-length) is refused."""
    import daggerfft as D
    from daggerfft import _lib as L
    a = O.fill(28, (8191, 2, 1), np.complex128)
    p = D.Plan(a.shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    assert p.passes()[0]["bluestein"] == 16384
    F = p.execute(_dev(a), D.DeviceArray(a.shape, np.complex128))
    assert O.rel_l2(F.to_host(), O.fftn(a)) <= TOL64
    with pytest.raises(D.DFFTError) as ei:
        D.Plan((8209, 2, 1), L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    assert ei.value.status == L.DFFT_ERR_UNSUPPORTED_SIZE
    D.Plan((4093 * 2, 2, 1), L.R2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    with pytest.raises(D.DFFTError) as ei:
        D.Plan((4099 * 2, 2, 1), L.R2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    assert ei.value.status == L.DFFT_ERR_UNSUPPORTED_SIZE


def test_autotune_covers_convolution_passes(gpu):
    """The autotuner times a convolution pass (Bluestein or Rader) over tile
    widths only, keeping its engine: tuning a shape with prime axes records
    an entry per axis, and the plan afterwards still runs convolution passes
    and matches the oracle."""
    import daggerfft as D
    from daggerfft import _lib as L
    shape = (11, 64, 13)
    try:
        D.tune_clear()
        D.autotune(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3), level=1)
        assert D.tune_entries() == 3  # the 64-point axis and both convolution passes
        p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
        assert [q["bluestein"] > 0 or q["rader"] > 0 for q in p.passes() if q["axis"] != 1] == [True, True]
        a = O.fill(27, shape, np.complex128)
        F = p.execute(_dev(a), D.DeviceArray(shape, np.complex128))
        assert O.rel_l2(F.to_host(), O.fftn(a)) <= TOL64
    finally:
        D.tune_clear()


@pytest.mark.parametrize("shape,prec", [((8191, 2, 1), 0), ((2, 8191, 1), 1), ((2, 8191, 1), 0), ((2, 1, 8191), 1)],
                         ids=["contig-f32", "strided-f64", "strided-f32", "plane-f64"])
def test_longest_bluestein_lines_every_layout(gpu, shape, prec):
    """The 16384-point convolution (n = 8191, prime) in both precisions,
    along the contiguous axis and both strided layouts."""
    import daggerfft as D
    from daggerfft import _lib as L
    dt = np.complex128 if prec else np.complex64
    a = O.fill(29, shape, dt)
    p = D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
    q = [x for x in p.passes() if x["L"] == 8191][0]
    assert q["bluestein"] == 16384 and q["kernel"].startswith("k_blue_fixed"), q
    F = p.execute(_dev(a), D.DeviceArray(shape, dt))
    assert O.rel_l2(F.to_host().astype(np.complex128), O.fftn(a.astype(np.complex128))) <= (TOL64 if prec else TOL32)


@pytest.mark.parametrize("n", [2 * 4093, 4093], ids=["packed-H4093", "full-4093"])
@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
def test_longest_real_bluestein_lines(gpu, n, prec):
    """rfft / irfft of the longest real lines with a prime factor above 7:
    packed R2C/C2R of half-length 4093 and R2C_FULL/C2R_FULL of length 4093,
    both on 8192-point convolutions."""
    import daggerfft as D
    from daggerfft import _lib as L
    rdt = np.float64 if prec else np.float32
    cdt = np.complex128 if prec else np.complex64
    tol = TOL64 if prec else TOL32
    shape = (n, 2, 1)
    x = O.fill(30, shape, rdt)
    pr = D.Plan(shape, L.R2C, -1, prec, D.Pencil(), (1, 2, 3))
    assert pr.passes()[0]["bluestein"] == 8192, pr.passes()[0]
    F = pr.execute(_dev(x), D.DeviceArray((n // 2 + 1, 2, 1), cdt))
    ref = O.rfftn(x.astype(np.float64))
    assert O.rel_l2(F.to_host().astype(np.complex128), ref) <= tol
    y = D.irfft(_dev(ref.astype(cdt)), n, decomp=D.Pencil(), dims=(1, 2, 3))
    assert O.rel_l2(y.to_host().astype(np.float64), x.astype(np.float64)) <= tol


def test_convolution_pass_tile_width_from_cache_and_tuner(gpu, tmp_path):
    """A convolution pass takes its tile width from the schedule cache along
    the contiguous axis too (Bluestein 131, Rader 257), gives the oracle's
    result with it, and the autotuner times tile widths of convolution
    passes (its entries are then in the cache)."""
    import math

    import daggerfft as D
    from daggerfft import _lib as L
    shipped = os.path.join(os.path.dirname(D.__file__), "..", "wisdom", "gfx950.txt")
    try:
        for n, var in ((131, 3), (257, 2)):
            shape = (n, 6, 4)
            sc = int(math.floor(math.log2(n * 24)))
            path = str(tmp_path / ("tc%d.txt" % n))
            with open(path, "w") as f:
                f.write("1/0/%d/1/%d - 0 4 0 %d 1\n" % (n, sc, var))
            D.tune_clear()
            D.tune_load(path)
            p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
            q = [x for x in p.passes() if x["axis"] == 0][0]
            assert q["TC"] == 4 and (q["rader"] if var == 2 else q["bluestein"]) > 0, q
            a = O.fill(40 + n, shape, np.complex128)
            F = p.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, np.complex128))
            assert O.rel_l2(F.to_host(), O.fftn(a)) <= 1e-12
        D.tune_clear()
        D.autotune((131, 8, 8), kind=L.C2C, sign=-1, precision=L.F64, level=1)
        assert D.tune_entries() >= 3
        p = D.Plan((131, 8, 8), L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
        a = O.fill(41, (131, 8, 8), np.complex128)
        F = p.execute(D.DeviceArray.from_host(a), D.DeviceArray((131, 8, 8), np.complex128))
        assert O.rel_l2(F.to_host(), O.fftn(a)) <= 1e-12
    finally:
        D.tune_clear()
        D.tune_load(shipped)
