"""Box gather / scatter / copy (pack_kernel!, unpack_kernel!,
local_copy_kernel!, src/fftgpu.jl:452-495) and the LDS-tiled axis permutation
on the GPU, bit-exact against the oracle's restatement of pack_data!,
unpack_data!, local_copy! (src/fft.jl:180-231)."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev(a):
    import daggerfft as D
    return D.DeviceArray.from_host(a)


def _s():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64, np.float32])
def test_pack_unpack_copy_bit_exact(gpu, dtype):
    import daggerfft as D
    from daggerfft import _lib as L
    src = O.fill(21, (37, 19, 11), dtype)
    lo, cnt, off = (3, 2, 4), (30, 15, 6), 5
    buf = np.zeros(off + int(np.prod(cnt)) + 3, dtype=dtype)
    O.pack(buf, src, lo, cnt, off)
    S = _dev(src)
    B = D.DeviceArray((buf.size,), dtype)
    B.tensor.zero_()
    es = np.dtype(dtype).itemsize
    L.check(L.lib().dfft_box_pack(ctypes.c_void_p(B.ptr), ctypes.c_void_p(S.ptr), L.i64(src.shape), L.i64(lo),
                                  L.i64(cnt), off, es, _s()))
    assert np.array_equal(B.to_host(), buf)
    dst = O.fill(22, (40, 20, 12), dtype)
    O.unpack(dst, buf, (1, 4, 5), cnt, off)
    Dd = _dev(O.fill(22, (40, 20, 12), dtype))
    L.check(L.lib().dfft_box_unpack(ctypes.c_void_p(Dd.ptr), ctypes.c_void_p(B.ptr), L.i64((40, 20, 12)),
                                    L.i64((1, 4, 5)), L.i64(cnt), off, es, _s()))
    assert np.array_equal(Dd.to_host(), dst)
    d2 = O.fill(23, (33, 17, 9), dtype)
    ref = d2.copy(order="F")
    O.local_copy(ref, (2, 1, 0), src, (5, 2, 3), (20, 12, 8))
    D2 = _dev(d2)
    L.check(L.lib().dfft_box_copy(ctypes.c_void_p(D2.ptr), L.i64((33, 17, 9)), L.i64((2, 1, 0)), ctypes.c_void_p(S.ptr),
                                  L.i64(src.shape), L.i64((5, 2, 3)), L.i64((20, 12, 8)), es, _s()))
    assert np.array_equal(D2.to_host(), ref)


@pytest.mark.parametrize("perm", [(1, 0, 2), (2, 0, 1), (1, 2, 0), (0, 2, 1), (2, 1, 0), (0, 1, 2)])
@pytest.mark.parametrize("dtype", [np.complex128, np.float64, np.float32])
def test_permute_bit_exact(gpu, perm, dtype):
    import daggerfft as D
    from daggerfft import _lib as L
    a = O.fill(31, (70, 33, 19), dtype)
    ref = O.permute(a, perm)
    A = _dev(a)
    out = D.DeviceArray(ref.shape, dtype)
    L.check(L.lib().dfft_permute(ctypes.c_void_p(out.ptr), ctypes.c_void_p(A.ptr), L.i64(a.shape), L.i32(perm),
                                 np.dtype(dtype).itemsize, _s()))
    assert np.array_equal(out.to_host(), ref)


def test_box_rejects_out_of_bounds(gpu):
    from daggerfft import _lib as L
    st = L.lib().dfft_box_pack(ctypes.c_void_p(16), ctypes.c_void_p(16), L.i64((4, 4, 4)), L.i64((2, 0, 0)),
                               L.i64((3, 1, 1)), 0, 8, None)
    assert st == L.DFFT_ERR_ARG
