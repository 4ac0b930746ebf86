"""R2C / C2R (rfft / irfft, README.md:38-44) and the R2R kinds (the DCT-II /
DCT-III pair of README.md:59-66 and every other kind of R2R_SUPPORTED_KINDS,
src/fft.jl:1-11) on the HIP path vs the oracle.
FP64 tolerance: relative L2 <= 1e-12."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def _dev(a):
    import daggerfft as D
    return D.DeviceArray.from_host(a)


@pytest.mark.parametrize("shape", [(16, 12, 10), (15, 6, 4), (2, 3, 3), (1, 5, 4), (64, 64, 64), (30, 7, 9)])
@pytest.mark.parametrize("decomp", ["pencil", "slab"])
def test_rfft_irfft_vs_oracle(gpu, shape, decomp):
    import daggerfft as D
    dec = D.Pencil() if decomp == "pencil" else D.Slab()
    a = O.fill(2, shape, np.float64)
    F = D.rfft(_dev(a), decomp=dec)
    ref = O.rfftn(a)
    assert F.shape == ref.shape
    e = O.rel_l2(F.to_host(), ref)
    print(shape, decomp, "rfft rel-L2", e)
    assert e <= TOL64
    B = D.irfft(F, shape[0], decomp=dec)
    assert B.shape == shape
    assert O.rel_l2(B.to_host(), O.irfftn(ref, shape[0])) <= TOL64
    assert O.rel_l2(B.to_host(), a) <= TOL64


@pytest.mark.parametrize("dims", [(2,), (3,), (2, 1), (3, 2, 1), (2, 3)])
def test_rfft_first_dim_strided(gpu, dims):
    import daggerfft as D
    shape = (12, 10, 14)
    a = O.fill(4, shape, np.float64)
    F = D.rfft(_dev(a), dims=dims)
    ref = O.rfftn(a, dims=dims)
    assert O.rel_l2(F.to_host(), ref) <= TOL64
    d = shape[dims[0] - 1]
    B = D.irfft(F, d, dims=dims)
    assert O.rel_l2(B.to_host(), a) <= TOL64


_R256 = {}


def _ref256():
    """Seed-2 Float64 256^3 input and its oracle rfft, computed once."""
    if not _R256:
        a = O.fill(2, (256, 256, 256), np.float64)
        _R256["a"] = a
        _R256["F"] = O.rfftn(a)
    return _R256["a"], _R256["F"]


@pytest.mark.parametrize("decomp", ["pencil", "slab"])
def test_rfft_256_cubed(gpu, decomp):
    """The 256^3 Float64 R2C/C2R configuration, both decompositions: every
    output element against the oracle (rfft), and irfft against the oracle's
    irfft of the same spectrum and against the input."""
    import daggerfft as D
    dec = D.Pencil() if decomp == "pencil" else D.Slab()
    shape = (256, 256, 256)
    a, ref = _ref256()
    A = D.rand(shape, np.float64, seed=2)
    F = D.rfft(A, decomp=dec)
    e = O.rel_l2(F.to_host(), ref)
    print("256^3", decomp, "rfft rel-L2", e)
    assert e <= TOL64
    B = D.irfft(F, 256, decomp=dec)
    e2 = O.rel_l2(B.to_host(), a)
    print("256^3", decomp, "irfft round trip rel-L2", e2)
    assert e2 <= TOL64
    # C2R alone, from the oracle's spectrum
    B2 = D.irfft(D.DeviceArray.from_host(ref), 256, decomp=dec)
    e3 = O.rel_l2(B2.to_host(), O.irfftn(ref, 256))
    print("256^3", decomp, "irfft rel-L2 vs oracle irfft", e3)
    assert e3 <= TOL64


def test_irfft_ignores_dc_nyquist_imag(gpu):
    import daggerfft as D
    shape = (8, 6)
    a = O.fill(3, shape, np.float64)
    ref = O.rfftn(a)
    bad = ref.copy()
    bad[0, :] += 0.25j
    bad[4, :] -= 0.5j
    B = D.irfft(_dev(np.asfortranarray(bad)), 8)
    assert O.rel_l2(B.to_host(), O.irfftn(bad, 8)) <= TOL64


def test_rfft_f32(gpu):
    import daggerfft as D
    shape = (64, 48, 30)
    A = D.rand(shape, np.float32, seed=12)
    a = O.fill(12, shape, np.float32).astype(np.float64)
    F = D.rfft(A)
    e = O.rel_l2(F.to_host().astype(np.complex128), O.rfftn(a))
    print("f32 rfft rel-L2 vs f64 oracle", e)
    assert e <= 1e-5
    B = D.irfft(F, 64)
    assert O.rel_l2(B.to_host().astype(np.float64), a) <= 1e-5


@pytest.mark.parametrize("shape,dims", [((256, 256), (1, 2)), ((128, 128, 128), (1, 2, 3)), ((7, 5, 3), (1, 2, 3)),
                                        ((1, 9, 4), (1, 2, 3)), ((30, 21, 10), (2, 3))])
def test_dct2_dct3_vs_oracle(gpu, shape, dims):
    import daggerfft as D
    a = O.fill(4, shape, np.float64)
    kinds = (D.REDFT10,) * len(dims)
    F = D.fft(_dev(a), decomp=D.Slab(), transforms=(D.R2R(kinds),), dims=dims)
    ref = O.r2rn(a, O.REDFT10, dims)
    e = O.rel_l2(F.to_host(), ref)
    print(shape, "REDFT10 rel-L2", e)
    assert e <= TOL64
    # FFTW-faithful REDFT01 (unnormalised) on the GPU vs the oracle
    from daggerfft import _lib as L
    p = D.Plan(shape, L.R2R, -1, L.F64, D.Slab(), dims, r2r_kinds=(D.REDFT01,) * len(dims))
    G = p.execute(F, F.similar())
    ref3 = O.r2rn(ref, O.REDFT01, dims)
    assert O.rel_l2(G.to_host(), ref3) <= TOL64
    # README ifft with REDFT01 reconstructs A (scaled by 1/prod(2n))
    B = D.ifft(F, decomp=D.Slab(), transforms=(D.R2R((D.REDFT01,) * len(dims)),), dims=dims)
    assert O.rel_l2(B.to_host(), a) <= TOL64


def test_mixed_r2r_kinds_per_dim(gpu):
    import daggerfft as D
    from daggerfft import _lib as L
    shape = (12, 10, 6)
    a = O.fill(8, shape, np.float64)
    p = D.Plan(shape, L.R2R, -1, L.F64, D.Pencil(), (1, 2, 3), r2r_kinds=(D.REDFT10, D.REDFT01, D.REDFT10))
    G = p.execute(_dev(a), D.DeviceArray(shape, np.float64))
    ref = O.r2rn(a, (O.REDFT10, O.REDFT01, O.REDFT10), (1, 2, 3))
    assert O.rel_l2(G.to_host(), ref) <= TOL64


def test_unsupported_r2r_kind_raises(gpu):
    """R2HC / HC2R are FFTW kinds outside the reference's R2R_SUPPORTED_KINDS
    (src/fft.jl:1-11); REDFT00 needs n > 1.  REDFT00 of n = 128 (an FFT of
    length 2 * 127) runs as a Bluestein pass."""
    import daggerfft as D
    from daggerfft import _lib as L
    with pytest.raises(D.DFFTError) as ei:
        D.Plan((8, 8), L.R2R, -1, L.F64, D.Pencil(), (1, 2), r2r_kinds=(D.R2HC, D.R2HC))
    assert ei.value.status == L.DFFT_ERR_UNSUPPORTED
    with pytest.raises(D.DFFTError) as ei:
        D.Plan((1, 8), L.R2R, -1, L.F64, D.Pencil(), (1, 2), r2r_kinds=(D.REDFT00, D.REDFT00))
    assert ei.value.status == L.DFFT_ERR_ARG
    a = O.fill(5, (128, 8), np.float64)
    G = D.fft(D.DeviceArray.from_host(a), decomp=D.Pencil(), dims=(1, 2), transforms=D.R2R((D.REDFT00, D.REDFT00)))
    assert O.rel_l2(G.to_host(), O.r2rn(a, (O.REDFT00, O.REDFT00), (1, 2))) <= TOL64


# Every kind of the reference's R2R_SUPPORTED_KINDS (src/fft.jl:1-11).  The
# complex FFT behind each kind has length n (DHT, DCT/DST-II/III), 2n
# (DCT/DST-IV), 2(n-1) (DCT-I) or 2(n+1) (DST-I); the shapes keep those
# lengths 7-smooth and include odd n and n = 1.
R2R_SHAPES = {
    "REDFT00": [(9, 6, 5), (129, 33, 17), (2, 3, 8)],
    "RODFT00": [(7, 5, 9), (127, 31, 15), (1, 2, 3)],
}
R2R_DEFAULT_SHAPES = [(12, 10, 7), (64, 48, 30), (9, 1, 5), (15, 14, 3)]
INVERSE = {"REDFT10": "REDFT01", "REDFT01": "REDFT10", "RODFT10": "RODFT01", "RODFT01": "RODFT10"}


@pytest.mark.parametrize("kind", ["DHT", "REDFT00", "REDFT01", "REDFT10", "REDFT11", "RODFT00", "RODFT01",
                                  "RODFT10", "RODFT11"])
@pytest.mark.parametrize("prec", [1, 0])
def test_every_r2r_kind_vs_oracle(gpu, kind, prec):
    """Each kind along every axis (Pencil, all three dims) and along two
    (Slab), against the oracle's long-double direct sums; then the README
    inverse (normalised by FFTW's logical size) reconstructs the input.
    Float64 rel-L2 <= 1e-12; Float32 against the FP64 oracle <= 1e-5 (stated)."""
    import daggerfft as D
    from daggerfft import _lib as L
    k = getattr(D, kind)
    ki = getattr(D, INVERSE.get(kind, kind))
    tol = 1e-12 if prec else 1e-5
    dt = np.float64 if prec else np.float32
    worst = 0.0
    for shape in R2R_SHAPES.get(kind, R2R_DEFAULT_SHAPES):
        a = O.fill(sum(shape), shape, dt)
        a64 = a.astype(np.float64)
        for decomp, dims in ((D.Pencil(), (1, 2, 3)), (D.Slab(), (2, 3))):
            print(kind, "f64" if prec else "f32", shape, dims, flush=True)
            p = D.Plan(shape, L.R2R, -1, prec, decomp, dims, r2r_kinds=(k,) * len(dims))
            G = p.execute(_dev(a), D.DeviceArray(shape, dt))
            ref = O.r2rn(a64, getattr(O, kind), dims)
            e = O.rel_l2(G.to_host().astype(np.float64), ref)
            worst = max(worst, e)
            assert e <= tol, (shape, dims, e)
            pi = D.Plan(shape, L.R2R, +1, prec, decomp, dims, r2r_kinds=(ki,) * len(dims),
                        flags=L.FLAG_NORMALIZE_R2R)
            B = pi.execute(G, D.DeviceArray(shape, dt))
            assert O.rel_l2(B.to_host().astype(np.float64), a64) <= tol, (shape, dims, "inverse")
    print(kind, "worst rel-L2 %.3e" % worst)
