"""Full-array parity at the headline size: 3D C2C ComplexF64 512x512x512
(seed 3), forward and inverse, Pencil and Slab, every output element against
the oracle (FFTW's definitions restated in oracle/dfft_oracle.c, per-axis
batched FFTs in the reference's stage order, src/fft.jl:651-776).  The oracle
runs one such transform in a few seconds on the host's cores.

Tolerance: relative L2 <= 1e-12 (the stated FP64 target); the largest
elementwise error relative to the largest output magnitude is printed too.

The commented check in the reference's benchmark compares the whole collected
result with FFTW the same way (benchmarks/fftmpi.jl:88-92)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL64 = 1e-12
SHAPE = (512, 512, 512)

_REF = {}


def _ref(direction):
    """The seed-3 input and the oracle's forward or inverse transform of it,
    computed once per session (2 GiB each)."""
    if "a" not in _REF:
        _REF["a"] = O.fill(3, SHAPE, np.complex128)
    if direction not in _REF:
        _REF[direction] = O.fftn(_REF["a"], inverse=(direction == "inv"))
    return _REF["a"], _REF[direction]


@pytest.mark.parametrize("direction", ["fwd", "inv"])
@pytest.mark.parametrize("decomp", ["pencil", "slab"])
def test_512_cubed_c2c_full_array(gpu, decomp, direction):
    import daggerfft as D
    dec = D.Pencil() if decomp == "pencil" else D.Slab()
    a, ref = _ref(direction)
    A = D.rand(SHAPE, np.complex128, seed=3)
    F = (D.fft if direction == "fwd" else D.ifft)(A, decomp=dec, dims=(1, 2, 3))
    y = F.to_host()
    del F, A
    e = O.rel_l2(y, ref)
    emax = float(np.max(np.abs(y - ref)) / np.max(np.abs(ref)))
    print("512^3 C2C F64 %s %s: rel-L2 %.3e, max |err| / max |ref| %.3e" % (decomp, direction, e, emax))
    assert e <= TOL64
