"""Full-size parity against the oracle without an oracle FFT: sampled output
bins of the 512^3 ComplexF64 C2C (Pencil and Slab), the 384x240x210 C2C and
the 256^3 Float64 R2C transforms, each bin computed on the host by direct
separable summation in long double (oracle.spot_bins, pinned against the
golden vectors).  Bins: DC, the Nyquist corners, unit bins along each axis,
the last bin and random bins from a fixed seed.  The sampled non-DC bins
agree to rel-L2 <= 1e-12 and the DC bin to 1e-12 relative."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bins(shape, nrand, seed, half0=False):
    n0 = shape[0] // 2 + 1 if half0 else shape[0]
    fixed = [(0, 0, 0), (n0 - 1, 0, 0), (0, shape[1] // 2, 0), (0, 0, shape[2] // 2), (1, 0, 0), (0, 1, 0),
             (0, 0, 1), (n0 - 1, shape[1] - 1, shape[2] - 1), (min(n0 - 1, shape[0] // 2), shape[1] // 2,
                                                                 shape[2] // 2)]
    rng = np.random.default_rng(seed)
    rand = [(int(rng.integers(n0)), int(rng.integers(shape[1])), int(rng.integers(shape[2]))) for _ in range(nrand)]
    return fixed + rand


def _check(got, ref):
    dc, rest = 0, slice(1, None)
    assert abs(got[dc] - ref[dc]) <= 1e-12 * abs(ref[dc])
    e = O.rel_l2(got[rest], ref[rest])
    print("spot bins: DC rel %.2e, others rel-L2 %.2e" % (abs(got[dc] - ref[dc]) / abs(ref[dc]), e))
    assert e <= 1e-12


@pytest.mark.parametrize("decomp", ["pencil", "slab"])
def test_spot_bins_512cubed_c2c(gpu, decomp):
    import daggerfft as D
    shape = (512, 512, 512)
    print("512^3 C2C", decomp, flush=True)
    X = D.rand(shape, np.complex128, seed=3)
    F = D.fft(X, decomp=D.Pencil() if decomp == "pencil" else D.Slab())
    bins = _bins(shape, 8, 3)
    Fh = F.to_host()
    got = np.array([Fh[b] for b in bins])
    del Fh, F, X
    a = O.fill(3, shape, np.complex128)
    _check(got, O.spot_bins(a, bins))


def test_spot_bins_384x240x210_c2c(gpu):
    import daggerfft as D
    shape = (384, 240, 210)
    X = D.rand(shape, np.complex128, seed=6)
    Fh = D.fft(X, decomp=D.Slab()).to_host()
    bins = _bins(shape, 12, 6)
    _check(np.array([Fh[b] for b in bins]), O.spot_bins(O.fill(6, shape, np.complex128), bins))


def test_spot_bins_256cubed_r2c(gpu):
    """rfft keeps bins 0..n0/2 of the first dim of the full DFT of the real input."""
    import daggerfft as D
    shape = (256, 256, 256)
    X = D.rand(shape, np.float64, seed=2)
    Fh = D.rfft(X).to_host()
    bins = _bins(shape, 12, 2, half0=True)
    a = O.fill(2, shape, np.float64).astype(np.complex128)
    _check(np.array([Fh[b] for b in bins]), O.spot_bins(a, bins))
