"""A seeded sweep of random transforms against the oracle: shapes of rank 1-3
with unit and odd edges, random dims subsets and orders, C2C forward and
inverse, R2C, C2R and every r2r kind of R2R_SUPPORTED_KINDS, both precisions,
Pencil and Slab.  The first set draws 7-smooth edges (Stockham passes), the
second any edge up to 140 and the third any edge up to 300 (primes and other
lengths with a factor above 7 run Bluestein passes).  ComplexF64/Float64 rel-L2 <= 1e-12; ComplexF32/Float32
against the FP64 oracle <= 2e-5.  The only refusal is REDFT00 of length 1,
which FFTW rejects too."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SMOOTH = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 18, 20, 21, 24, 25, 27, 28, 30, 32, 35, 36, 40, 42, 45,
          48, 49, 50, 54, 56, 60, 63, 64, 70, 72, 75, 80, 81, 84, 90, 96, 98, 100, 105, 108, 112, 120, 125, 126, 128]
KINDS = ["DHT", "REDFT00", "REDFT01", "REDFT10", "REDFT11", "RODFT00", "RODFT01", "RODFT10", "RODFT11"]


def _smooth(m):
    for p in (2, 3, 5, 7):
        while m % p == 0:
            m //= p
    return m == 1


ANY = list(range(1, 141))
ANY300 = list(range(1, 301))


def _r2r_ok(kind, n):
    """FFTW's REDFT00 needs n > 1; every other kind and length runs."""
    return not (kind == "REDFT00" and n < 2)


def _cases(count=200, seed=2024, sizes=SMOOTH, first=0):
    rng = np.random.default_rng(seed)
    out = []
    for c in range(first, first + count):
        rank = int(rng.integers(1, 4))
        shape = tuple(int(rng.choice(sizes)) for _ in range(rank))
        while np.prod(shape) > 400_000:  # keeps the host oracle under a second
            shape = tuple(int(rng.choice(sizes)) for _ in range(rank))
        nd = int(rng.integers(1, rank + 1))
        dims = tuple(int(d) + 1 for d in rng.permutation(rank)[:nd])
        op = str(rng.choice(["c2c", "ic2c", "r2c", "c2r", "r2r"]))
        prec = int(rng.integers(0, 2))
        decomp = "slab" if (rank == 3 and nd == 3 and rng.random() < 0.4) else "pencil"
        kinds = tuple(str(rng.choice(KINDS)) for _ in range(nd)) if op == "r2r" else None
        if op in ("r2c", "c2r") and shape[dims[0] - 1] < 2:
            op = "c2c"
        out.append((c, shape, dims, op, prec, decomp, kinds))
    return out


@pytest.mark.parametrize("case", _cases() + _cases(120, 4049, ANY, 200) + _cases(150, 777, ANY300, 320), ids=lambda c: "%d-%s-%s-%s-f%d-%s" % (
    c[0], "x".join(map(str, c[1])), "".join(map(str, c[2])), c[3], 64 if c[4] else 32, c[5]))
def test_sweep_against_oracle(gpu, case):
    import daggerfft as D
    from daggerfft import _lib as L
    c, shape, dims, op, prec, decomp, kinds = case
    dec = D.Slab() if decomp == "slab" else D.Pencil()
    tol = 1e-12 if prec else 2e-5
    cdt = np.complex128 if prec else np.complex64
    rdt = np.float64 if prec else np.float32
    if op in ("c2c", "ic2c"):
        a = O.fill(100 + c, shape, cdt)
        X = D.DeviceArray.from_host(a)
        Y = D.fft(X, decomp=dec, dims=dims) if op == "c2c" else D.ifft(X, decomp=dec, dims=dims)
        ref = O.fftn(a.astype(np.complex128), dims, inverse=(op == "ic2c"))
    elif op == "r2c":
        a = O.fill(100 + c, shape, rdt)
        Y = D.rfft(D.DeviceArray.from_host(a), decomp=dec, dims=dims)
        ref = O.rfftn(a.astype(np.float64), dims)
    elif op == "c2r":
        n = shape[dims[0] - 1]
        x = O.fill(100 + c, shape, np.float64)
        F = O.rfftn(x, dims)
        Y = D.irfft(D.DeviceArray.from_host(F.astype(cdt)), n, decomp=dec, dims=dims)
        ref = x
    else:
        a = O.fill(100 + c, shape, rdt)
        ok = all(_r2r_ok(k, shape[d - 1]) for k, d in zip(kinds, dims))
        if not ok:
            with pytest.raises(L.DFFTError):
                D.fft(D.DeviceArray.from_host(a), decomp=dec, dims=dims,
                      transforms=D.R2R(tuple(getattr(D, k) for k in kinds)))
            return
        Y = D.fft(D.DeviceArray.from_host(a), decomp=dec, dims=dims,
                  transforms=D.R2R(tuple(getattr(D, k) for k in kinds)))
        ref = O.r2rn(a.astype(np.float64), tuple(getattr(O, k) for k in kinds), dims)
    got = Y.to_host()
    got = got.astype(np.complex128 if np.iscomplexobj(got) else np.float64)
    e = O.rel_l2(got, ref) if np.linalg.norm(ref) > 0 else float(np.abs(got).max())
    assert e <= tol, (case, e)
