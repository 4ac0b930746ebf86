"""Chunked arrays on the GPU (the reference's Blocks-distributed DArray with
every chunk on one device): distribute/collect and the box-copy
redistribution are bit-exact against the oracle's restatement of
local_copy!/redistribution (src/fft.jl:216-231, 453-462), and the Pencil
and Slab stage drivers over chunked arrays (src/fft.jl:651-776) match the
oracle's chunked pipeline: ComplexF64 rel-L2 <= 1e-12."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPE = (24, 20, 18)
CH = 8   # blocks of 8 leave uneven edge chunks in every dim


def _host_chunks(D):
    return [c.to_host() for c in D.chunks]


@pytest.mark.parametrize("dtype", [np.complex128, np.float32])
def test_distribute_collect_redistribute_bit_exact(gpu, dtype):
    import daggerfft as D
    from daggerfft import chunked as C
    a = O.fill(41, SHAPE, dtype)
    X = C.DArray.distribute(D.DeviceArray.from_host(a), (SHAPE[0], CH, CH))
    src = O.Blocked.distribute(a, (SHAPE[0], CH, CH))
    for g, r in zip(_host_chunks(X), src.chunks):
        assert np.array_equal(g, r)
    Y = C.DArray(SHAPE, (CH, SHAPE[1], CH), dtype)
    ws = D.create_workspace(X, Y)
    C.redistribute(Y, X, ws)
    ref = O.Blocked(SHAPE, (CH, SHAPE[1], CH), dtype)
    O.redistribute(ref, src)
    for g, r in zip(_host_chunks(Y), ref.chunks):
        assert np.array_equal(g, r)
    assert np.array_equal(Y.collect().to_host(), a)


def test_chunked_pencil_fft_ifft(gpu):
    import daggerfft as D
    a = O.fill(42, SHAPE, np.complex128)
    bl = {"x": (SHAPE[0], CH, CH), "y": (CH, SHAPE[1], CH), "z": (CH, CH, SHAPE[2])}
    A = D.DArray.distribute(D.DeviceArray.from_host(a), bl["x"])
    B = D.DArray(SHAPE, bl["y"], np.complex128)
    Cz = D.DArray(SHAPE, bl["z"], np.complex128)
    ws_ab, ws_bc = D.create_workspace(A, B), D.create_workspace(B, Cz)
    out = D.fft_(Cz, A, B, ws_ab, ws_bc, (D.FFT_(), D.FFT_(), D.FFT_()), (1, 2, 3))
    ref = O.fft_pencil(a, chunk=CH)
    F = out.collect().to_host()
    assert O.rel_l2(F, ref) <= 1e-12
    assert O.rel_l2(F, O.fftn(a)) <= 1e-12
    # inverse: z-pencils in, x-pencils out (src/fft.jl:691-724)
    Ai = D.DArray.distribute(D.DeviceArray.from_host(F), bl["z"])
    Bi = D.DArray(SHAPE, bl["y"], np.complex128)
    Ci = D.DArray(SHAPE, bl["x"], np.complex128)
    back = D.ifft_(Ci, Ai, Bi, D.create_workspace(Ai, Bi), D.create_workspace(Bi, Ci),
                   (D.IFFT_(), D.IFFT_(), D.IFFT_()), (1, 2, 3))
    assert O.rel_l2(back.collect().to_host(), O.ifft_pencil(F, chunk=CH)) <= 1e-12
    assert O.rel_l2(back.collect().to_host(), a) <= 1e-12


def test_chunked_pencil_prime_edges(gpu):
    """Chunked arrays with prime edges (Bluestein passes inside each chunk's
    apply_fft!) and uneven blocks."""
    import daggerfft as D
    shape, ch = (23, 19, 13), 7
    a = O.fill(44, shape, np.complex128)
    A = D.DArray.distribute(D.DeviceArray.from_host(a), (shape[0], ch, ch))
    B = D.DArray(shape, (ch, shape[1], ch), np.complex128)
    Cz = D.DArray(shape, (ch, ch, shape[2]), np.complex128)
    out = D.fft_(Cz, A, B, D.create_workspace(A, B), D.create_workspace(B, Cz), (D.FFT_(), D.FFT_(), D.FFT_()),
                 (1, 2, 3))
    F = out.collect().to_host()
    assert O.rel_l2(F, O.fft_pencil(a, chunk=ch)) <= 1e-12
    assert O.rel_l2(F, O.fftn(a)) <= 1e-12


def test_chunked_slab_fft_ifft(gpu):
    import daggerfft as D
    a = O.fill(43, SHAPE, np.complex128)
    A = D.DArray.distribute(D.DeviceArray.from_host(a), (SHAPE[0], SHAPE[1], CH))
    B = D.DArray(SHAPE, (CH, SHAPE[1], SHAPE[2]), np.complex128)
    out = D.fft_(B, A, D.create_workspace(A, B), (D.FFT_(), D.FFT_(), D.FFT_()), (1, 2, 3), D.Slab())
    F = out.collect().to_host()
    assert O.rel_l2(F, O.fft_slab(a, chunk=CH)) <= 1e-12
    # ifft!(A, B, ws_BA, ..., Slab()): dims[3] on B, then (dims[1], dims[2]) on A
    A2 = D.DArray(SHAPE, (SHAPE[0], SHAPE[1], CH), np.complex128)
    back = D.ifft_(A2, out, D.create_workspace(out, A2), (D.IFFT_(), D.IFFT_(), D.IFFT_()), (1, 2, 3), D.Slab())
    assert O.rel_l2(back.collect().to_host(), O.ifft_slab(F, chunk=CH)) <= 1e-12
    assert O.rel_l2(back.collect().to_host(), a) <= 1e-12


def test_chunked_pencil_r2r(gpu):
    """R2R stages (DCT-II on every dim) through the chunked Pencil driver
    equal the oracle's separable REDFT10."""
    import daggerfft as D
    a = O.fill(44, SHAPE, np.float64)
    A = D.DArray.distribute(D.DeviceArray.from_host(a), (SHAPE[0], CH, CH))
    B = D.DArray(SHAPE, (CH, SHAPE[1], CH), np.float64)
    Cz = D.DArray(SHAPE, (CH, CH, SHAPE[2]), np.float64)
    t = D.R2R((D.REDFT10,))
    out = D.fft_(Cz, A, B, D.create_workspace(A, B), D.create_workspace(B, Cz), (t, t, t), (1, 2, 3))
    ref = O.r2rn(a, (O.REDFT10,) * 3)
    assert O.rel_l2(out.collect().to_host(), ref) <= 1e-12
