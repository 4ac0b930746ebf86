"""Cross-check against the platform's FFT library (rocFFT through torch.fft,
hipFFT) at the configuration sizes, including the full 512^3 ComplexF64
block: 3D C2C forward and inverse, R2C and C2R, both decompositions, both
precisions.  This is a second, independent implementation beside the oracle
(tests/test_gpu_fullsize.py compares the same 512^3 block with the oracle).  The library under test runs
through its C-ABI; torch.fft is only the comparison (an independent
implementation, so agreement pins sign, normalisation, axis order and the
half-spectrum layout at full size).

Tolerances: ComplexF64 rel-L2 <= 1e-12; ComplexF32 <= 2e-6 (both sides round
in FP32, the worst value is printed).

A flat column-major (n1, n2, n3) block read as a row-major (n3, n2, n1)
tensor is the same 3D array, so torch.fft over all dims is the same DFT and
rfftn halves torch's last dim = this library's first dim."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-12, np.complex64: 2e-6, np.float64: 1e-12, np.float32: 2e-6}


def _rel(a, b):
    import torch
    return float(torch.linalg.vector_norm(a.reshape(-1) - b.reshape(-1)) / torch.linalg.vector_norm(b.reshape(-1)))


@pytest.mark.parametrize("shape,dt,decomp", [
    ((512, 512, 512), np.complex128, "pencil"),
    ((512, 512, 512), np.complex128, "slab"),
    ((384, 240, 210), np.complex128, "slab"),
    ((128, 128, 128), np.complex128, "pencil"),
    ((256, 192, 160), np.complex64, "pencil"),
    ((210, 98, 486), np.complex64, "slab"),
])
def test_c2c_vs_rocfft(gpu, shape, dt, decomp):
    import torch
    import daggerfft as D
    dec = D.Pencil() if decomp == "pencil" else D.Slab()
    print("C2C", shape, np.dtype(dt).name, decomp, flush=True)
    A = D.rand(shape, dt, seed=11)
    F = D.fft(A, decomp=dec)
    T = A.tensor.view(shape[::-1])
    ref = torch.fft.fftn(T)
    e = _rel(F.tensor, ref)
    print("  forward rel-L2 vs rocFFT %.3e" % e, flush=True)
    assert e <= TOL[dt], e
    del ref
    B = D.ifft(F, decomp=dec)
    refi = torch.fft.ifftn(F.tensor.view(shape[::-1]))
    e2 = _rel(B.tensor, refi)
    print("  inverse rel-L2 vs rocFFT %.3e" % e2, flush=True)
    assert e2 <= TOL[dt], e2
    del A, F, B, refi, T
    torch.cuda.empty_cache()


@pytest.mark.parametrize("shape,dt,decomp", [
    ((256, 256, 256), np.float64, "pencil"),
    ((256, 256, 256), np.float64, "slab"),
    ((240, 210, 98), np.float32, "pencil"),
])
def test_r2c_c2r_vs_rocfft(gpu, shape, dt, decomp):
    import torch
    import daggerfft as D
    dec = D.Pencil() if decomp == "pencil" else D.Slab()
    print("R2C", shape, np.dtype(dt).name, decomp, flush=True)
    A = D.rand(shape, dt, seed=12)
    F = D.rfft(A, decomp=dec)
    ref = torch.fft.rfftn(A.tensor.view(shape[::-1]))
    assert tuple(ref.shape[::-1]) == F.shape == (shape[0] // 2 + 1,) + shape[1:]
    e = _rel(F.tensor, ref)
    print("  rfft rel-L2 vs rocFFT %.3e" % e, flush=True)
    assert e <= TOL[dt], e
    B = D.irfft(F, shape[0], decomp=dec)
    refi = torch.fft.irfftn(ref, s=shape[::-1])
    e2 = _rel(B.tensor, refi)
    print("  irfft rel-L2 vs rocFFT %.3e" % e2, flush=True)
    assert e2 <= TOL[dt], e2
    del A, F, B, ref, refi
    torch.cuda.empty_cache()
