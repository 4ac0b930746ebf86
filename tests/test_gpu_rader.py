"""Rader passes (prime line lengths n with 7-smooth n-1; Rader 1968) against
the oracle: every prime of the compile-time catalog (11 .. 8101) along the
contiguous axis and both strided layouts, in both precisions, forward and
inverse; full 3D transforms of 257^3 and 251^3 ComplexF64 (every element);
in place; against the Bluestein engines forced through the schedule cache on
the same lengths.

Tolerances: ComplexF64 rel-L2 <= 1e-12; ComplexF32 against the FP64 oracle
<= 1e-5 (the worst is printed)."""
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL64, TOL32 = 1e-12, 1e-5
SHIPPED = os.path.join(ROOT, "daggerfft.jl_amd", "wisdom", "gfx950.txt")


def _smooth(m):
    for p in (2, 3, 5, 7):
        while m % p == 0:
            m //= p
    return m == 1


def _prime(n):
    return n > 1 and all(n % q for q in range(2, int(n ** 0.5) + 1))


RADER_PRIMES = [n for n in range(11, 8102) if _prime(n) and _smooth(n - 1)]


def _dev(a):
    import daggerfft as D
    return D.DeviceArray.from_host(a)


@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
@pytest.mark.parametrize("chunk", range(4))
def test_every_rader_prime_every_layout(gpu, tmp_path, prec, chunk):
    """Every Rader kernel, forced through the schedule cache (variant 2: the
    planner takes Rader by default only where its FFT has no more stages
    than the Bluestein alternative)."""
    import math

    import daggerfft as D
    from daggerfft import _lib as L
    assert len(RADER_PRIMES) == 91
    dt = np.complex128 if prec else np.complex64
    tol = TOL64 if prec else TOL32
    worst = 0.0
    path = str(tmp_path / "rader.txt")
    try:
        for n in RADER_PRIMES[chunk::4]:
            for shape, ax, layout in (((n, 3, 2), 0, 1), ((3, n, 2), 1, 2), ((2, 3, n), 2, 0)):
                with open(path, "w") as f:
                    f.write("%d/0/%d/%d/%d - 0 0 0 2 1\n" % (prec, n, layout, int(math.floor(math.log2(6 * n)))))
                D.tune_clear()
                D.tune_load(path)
                a = O.fill(n + 7 * ax, shape, dt)
                a64 = a.astype(np.complex128)
                sign = -1 if (n + ax) % 2 else +1
                p = D.Plan(shape, L.C2C, sign, prec, D.Pencil(), (1, 2, 3))
                q = [x for x in p.passes() if x["axis"] == ax][0]
                assert q["rader"] == n - 1 and q["kernel"].startswith("k_rader_fixed"), (n, ax, q)
                F = p.execute(_dev(a), D.DeviceArray(shape, dt))
                e = O.rel_l2(F.to_host().astype(np.complex128), O.fftn(a64, inverse=sign > 0))
                worst = max(worst, e)
                assert e <= tol, (n, shape, sign, e)
    finally:
        D.tune_clear()
        D.tune_load(SHIPPED)
    print("chunk %d %s: worst rel-L2 %.2e" % (chunk, "f64" if prec else "f32", worst))


def test_default_choice_between_rader_and_bluestein(gpu):
    """Without a cache entry a prime length runs Rader exactly when its
    (n-1)-point FFT has no more stages than the compile-time Bluestein FFT
    the planner would pick: 257 (8 x 8 x 4 or 16 x 16) takes Rader."""
    import daggerfft as D
    from daggerfft import _lib as L
    try:
        D.tune_clear()
        for prec in (1, 0):
            for shape in ((257, 4, 2), (4, 257, 2), (4, 2, 257)):
                q = [x for x in D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3)).passes() if x["L"] == 257][0]
                assert q["rader"] == 256, q
            for n in RADER_PRIMES:
                q = D.Plan((n, 2, 2), L.C2C, -1, prec, D.Pencil(), (1, 2, 3)).passes()[0]
                assert (q["rader"] == n - 1) != (q["bluestein"] >= 2 * n - 1), q
    finally:
        D.tune_clear()
        D.tune_load(SHIPPED)


_REF = {}


@pytest.mark.parametrize("n", [257, 251])
@pytest.mark.parametrize("direction", ["fwd", "inv"])
def test_rader_cubes_full_array(gpu, n, direction):
    """257^3 (Rader over 256) and 251^3 (Rader over 250 or Bluestein over
    512, whichever the planner picks) ComplexF64, Pencil, every element."""
    import daggerfft as D
    shape = (n, n, n)
    key = (n, direction)
    if key not in _REF:
        a = O.fill(5, shape, np.complex128)
        _REF[key] = O.fftn(a, inverse=direction == "inv")
    A = D.rand(shape, np.complex128, seed=5)
    F = (D.fft if direction == "fwd" else D.ifft)(A)
    p = D.plan_fft(A) if direction == "fwd" else D.plan_ifft(A)
    print([(q["rader"], q["bluestein"], q["radices"]) for q in p.passes()])
    e = O.rel_l2(F.to_host(), _REF[key])
    print("%d^3 %s rel-L2 %.2e" % (n, direction, e))
    assert e <= TOL64


def test_rader_in_place_and_dims(gpu):
    import daggerfft as D
    from daggerfft import _lib as L
    shape = (97, 64, 193)
    a = O.fill(9, shape, np.complex128)
    A = _dev(a)
    p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    p.execute(A, A)
    assert O.rel_l2(A.to_host(), O.fftn(a)) <= TOL64
    for dims in ((1,), (3,), (3, 1), (2, 3)):
        F = D.fft(_dev(a), dims=dims)
        assert O.rel_l2(F.to_host(), O.fftn(a, dims=dims)) <= TOL64, dims
        B = D.ifft(F, dims=dims)
        assert O.rel_l2(B.to_host(), a) <= TOL64, dims


@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
def test_rader_against_forced_bluestein(gpu, tmp_path, prec):
    """The same 257- and 1009-point lines through Rader and through the
    compile-time and runtime Bluestein engines (schedule-cache variants 2, 3
    and 0): all three agree with the oracle and with each other."""
    import math

    import daggerfft as D
    from daggerfft import _lib as L
    dt = np.complex128 if prec else np.complex64
    tol = TOL64 if prec else TOL32
    path = str(tmp_path / "force.txt")
    try:
        for shape, ax, layout in (((257, 6, 4), 0, 1), ((6, 1009, 4), 1, 2), ((6, 4, 257), 2, 0)):
            n = shape[ax]
            a = O.fill(11, shape, dt)
            ref = O.fftn(a.astype(np.complex128))
            sc = int(math.floor(math.log2(np.prod(shape))))
            got = {}
            for variant in (2, 3, 0):
                D.tune_clear()
                with open(path, "w") as f:
                    f.write("%d/0/%d/%d/%d - 0 0 0 %d 1\n" % (prec, n, layout, sc, variant))
                D.tune_load(path)
                p = D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
                q = [x for x in p.passes() if x["axis"] == ax][0]
                kind = q["kernel"].split("<")[0]
                assert kind == {2: "k_rader_fixed", 3: "k_blue_fixed", 0: "k_axis_blue"}[variant], q
                F = p.execute(_dev(a), D.DeviceArray(shape, dt)).to_host().astype(np.complex128)
                assert O.rel_l2(F, ref) <= tol, (shape, variant)
                got[variant] = F
            assert O.rel_l2(got[2], got[3]) <= 2 * tol
    finally:
        D.tune_clear()
        D.tune_load(SHIPPED)
