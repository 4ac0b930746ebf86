"""Arrays above 2^31 elements: every global offset in the kernels is 64-bit
(tile bases, strides, the fill generator's counter).  A 2048 x 1050 x 1024
ComplexF64 C2C (2.2e9 elements, 35 GB per array) is checked at sampled bins
against the oracle's long-double sums (every bin sums every input element,
and the last bin is stored at the highest offset) and by its round trip; the
Float64 R2C of the same shape likewise.  Tolerances: rel-L2 <= 1e-12."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
SHAPE = (2048, 1050, 1024)


def _flat(shape, b):
    return b[0] + shape[0] * (b[1] + shape[1] * b[2])


def _bins(shape, nrand, seed):
    rng = np.random.default_rng(seed)
    fixed = [(0, 0, 0), (shape[0] - 1, shape[1] - 1, shape[2] - 1), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    return fixed + [tuple(int(rng.integers(n)) for n in shape) for _ in range(nrand)]


def _gather(arr, bins):
    import torch
    idx = torch.tensor([_flat(arr.shape, b) for b in bins], device=arr.tensor.device)
    return arr.tensor[idx].cpu().numpy().astype(np.complex128)


def _rel(a, b):
    import torch
    return (torch.linalg.vector_norm(a.tensor - b.tensor) / torch.linalg.vector_norm(b.tensor)).item()


def _check(got, ref):
    e_dc = abs(got[0] - ref[0]) / abs(ref[0])
    e = O.rel_l2(got[1:], ref[1:])
    print("spot bins: DC rel %.2e, others rel-L2 %.2e" % (e_dc, e), flush=True)
    assert e_dc <= 1e-12 and e <= 1e-12


def test_same_line_lengths_small(gpu):
    """The passes of the large cases (2048-point contiguous lines, 1050-point
    strided lines, and the R2C/C2R pair) at a size the full oracle checks."""
    import daggerfft as D
    shape = SHAPE[:2] + (3,)
    a = O.fill(8, shape, np.complex128)
    F = D.fft(D.DeviceArray.from_host(a), decomp=D.Pencil())
    assert O.rel_l2(F.to_host(), O.fftn(a)) <= 1e-12
    r = O.fill(8, shape, np.float64)
    G = D.rfft(D.DeviceArray.from_host(r), decomp=D.Pencil())
    assert O.rel_l2(G.to_host(), O.rfftn(r)) <= 1e-12
    assert O.rel_l2(D.irfft(G, shape[0], decomp=D.Pencil()).to_host(), r) <= 1e-12


@pytest.mark.timeout(900)
def test_c2c_above_2_31_elements(gpu):
    import torch
    import daggerfft as D
    assert np.prod(SHAPE) > 2 ** 31
    print("C2C", SHAPE, flush=True)
    X = D.rand(SHAPE, np.complex128, seed=9)
    F = D.fft(X, decomp=D.Pencil())
    bins = _bins(SHAPE, 5, 9)
    got = _gather(F, bins)
    R = D.ifft(F, decomp=D.Pencil())
    del F
    e_rt = _rel(R, X)
    del R, X
    torch.cuda.empty_cache()
    print("round trip rel-L2 %.2e" % e_rt, flush=True)
    assert e_rt <= 1e-12
    a = O.fill(9, SHAPE, np.complex128)
    ref = O.spot_bins(a, bins)
    del a
    _check(got, ref)


@pytest.mark.timeout(900)
def test_r2c_above_2_31_elements(gpu):
    import torch
    import daggerfft as D
    print("R2C", SHAPE, flush=True)
    x = D.rand(SHAPE, np.float64, seed=10)
    F = D.rfft(x, decomp=D.Pencil())
    half = (SHAPE[0] // 2 + 1,) + SHAPE[1:]
    assert F.shape == half
    bins = _bins(half, 3, 10)
    got = _gather(F, bins)
    y = D.irfft(F, SHAPE[0], decomp=D.Pencil())
    del F
    e_rt = _rel(y, x)
    del y, x
    torch.cuda.empty_cache()
    print("round trip rel-L2 %.2e" % e_rt, flush=True)
    assert e_rt <= 1e-12
    a = O.fill(10, SHAPE, np.float64).astype(np.complex128)
    ref = O.spot_bins(a, bins)
    del a
    _check(got, ref)
