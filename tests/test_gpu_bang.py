"""The exported bang API over single-chunk device arrays (src/DaggerFFT.jl:32-37):
fft!/ifft! with a per-stage transforms tuple.  Stage i applies transforms[i]
along dims[i] (Pencil, src/fft.jl:651-724); Slab applies transforms[1] over
(dims[1], dims[2]) and transforms[3] over dims[3] (src/fft.jl:727-776).
Mixed tuples such as (FFT!(), IFFT!(), FFT!()) are checked against the
oracle's stage drivers; RFFT!/IRFFT! stages raise.

Tolerance: ComplexF64 relative L2 <= 1e-12."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL64 = 1e-12

_NAMES = {"F": "FFT!", "I": "IFFT!"}


def _tags(code):
    import daggerfft as D
    return tuple(D.FFT_() if c == "F" else D.IFFT_() for c in code)


@pytest.mark.parametrize("code", ["FFF", "III", "FIF", "IFI", "FFI", "IIF"])
@pytest.mark.parametrize("dims", [(1, 2, 3), (3, 1, 2)])
def test_pencil_bang_per_stage(gpu, code, dims):
    import daggerfft as D
    shape = (24, 20, 14)
    a = O.fill(21, shape, np.complex128)
    A = D.DeviceArray.from_host(a)
    B, C = A.similar(), A.similar()
    ws_ab, ws_bc = D.create_workspace(A, B), D.create_workspace(B, C)
    ts = _tags(code)
    names = tuple(_NAMES[c] for c in code)
    out = D.fft_(C, A, B, ws_ab, ws_bc, ts, dims)
    ref = O.fft_pencil(a, names, dims)
    e = O.rel_l2(out.to_host(), ref)
    print("fft! Pencil", code, dims, "rel-L2", e)
    assert e <= TOL64
    A2 = D.DeviceArray.from_host(a)
    C2 = A2.similar()
    out2 = D.ifft_(C2, A2, A2.similar(), ws_ab, ws_bc, ts, dims)
    ref2 = O.ifft_pencil(a, names, dims)
    e2 = O.rel_l2(out2.to_host(), ref2)
    print("ifft! Pencil", code, dims, "rel-L2", e2)
    assert e2 <= TOL64


@pytest.mark.parametrize("code", ["FFF", "IFI", "FIF", "IIF", "FFI"])
def test_slab_bang_per_stage(gpu, code):
    """Slab: transforms[2] is ignored, so 'FIF' equals 'FFF'."""
    import daggerfft as D
    shape = (20, 18, 12)
    a = O.fill(22, shape, np.complex128)
    A = D.DeviceArray.from_host(a)
    B = A.similar()
    ws = D.create_workspace(A, B)
    ts = _tags(code)
    names = tuple(_NAMES[c] for c in code)
    out = D.fft_(B, A, ws, ts, (1, 2, 3), D.Slab())
    e = O.rel_l2(out.to_host(), O.fft_slab(a, names, (1, 2, 3)))
    print("fft! Slab", code, "rel-L2", e)
    assert e <= TOL64
    Bd = D.DeviceArray.from_host(a)
    Ad = Bd.similar()
    out2 = D.ifft_(Ad, Bd, ws, ts, (1, 2, 3), D.Slab())
    e2 = O.rel_l2(out2.to_host(), O.ifft_slab(a, names, (1, 2, 3)))
    print("ifft! Slab", code, "rel-L2", e2)
    assert e2 <= TOL64


def test_bang_mixed_tuple_is_not_uniform(gpu):
    """The mixed tuple really differs from the all-forward one (so the test
    above would catch a driver that looked only at transforms[1])."""
    import daggerfft as D
    shape = (16, 12, 10)
    a = O.fill(23, shape, np.complex128)
    A = D.DeviceArray.from_host(a)
    ws = D.create_workspace(A, A)
    mixed = D.fft_(A.similar(), A, A.similar(), ws, ws, _tags("FIF"), (1, 2, 3)).to_host()
    uniform = D.fft_(A.similar(), A, A.similar(), ws, ws, _tags("FFF"), (1, 2, 3)).to_host()
    assert O.rel_l2(mixed, uniform) > 0.1


@pytest.mark.parametrize("bad", ["RFFT_", "IRFFT_", "RFFT"])
def test_bang_rejects_real_tags(gpu, bad):
    import daggerfft as D
    shape = (8, 6, 4)
    A = D.rand(shape, np.complex128, seed=1)
    ws = D.create_workspace(A, A)
    t = getattr(D, bad)()
    with pytest.raises(ValueError):
        D.fft_(A.similar(), A, A.similar(), ws, ws, (D.FFT_(), t, D.FFT_()), (1, 2, 3))
    with pytest.raises(ValueError):
        D.ifft_(A.similar(), A, A.similar(), ws, ws, (t, t, t), (1, 2, 3))


def test_bang_r2r_per_stage(gpu):
    """R2R stages with a different kind per stage: DCT-II, DCT-III, DST-II."""
    import daggerfft as D
    shape = (16, 10, 12)
    a = O.fill(24, shape, np.float64)
    A = D.DeviceArray.from_host(a)
    ws = D.create_workspace(A, A)
    ts = (D.R2R((D.REDFT10,)), D.R2R((D.REDFT01,)), D.R2R((D.RODFT10,)))
    out = D.fft_(A.similar(), A, A.similar(), ws, ws, ts, (1, 2, 3))
    ref = O.r2rn(a, (D.REDFT10, D.REDFT01, D.RODFT10), (1, 2, 3))
    e = O.rel_l2(out.to_host(), ref)
    print("fft! R2R per stage rel-L2", e)
    assert e <= TOL64
