"""Every kernel variant of an axis pass gives the oracle's result:
  variant 0: runtime radix schedule (any length),
  variant 1: compile-time schedule of the length (the default where built),
  variant 2: compile-time schedule, 16 values per thread, real and imaginary
             parts exchanged through half the LDS (strided passes),
  variant 3: compile-time schedule, 16 values per thread (mixed radix;
             every length is covered in test_gpu_catalog.py).
Variants are forced through the schedule cache (dfft_tune_load)."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cache(tmp_path, shape, prec, variant, tc):
    """schedule-cache lines forcing `variant` for the C2C passes of `shape`"""
    N = int(np.prod(shape))
    sc = int(math.floor(math.log2(N)))
    lines = []
    for L in set(shape):
        for layout in (0, 1, 2):
            v = variant if (layout != 1 or variant < 2) else 1
            lines.append("%d/0/%d/%d/%d - 0 %d 0 %d" % (prec, L, layout, sc, 1 if layout == 1 else tc, v))
    p = os.path.join(str(tmp_path), "tune.txt")
    with open(p, "w") as f:
        f.write("\n".join(lines) + "\n")
    return p


@pytest.mark.parametrize("variant,tc", [(0, 16), (1, 8), (2, 16), (2, 32)])
@pytest.mark.parametrize("shape", [(64, 128, 256), (512, 64, 32), (1024, 16, 8)])
def test_c2c_variants_match_oracle(gpu, tmp_path, variant, tc, shape):
    import daggerfft as D
    from daggerfft import _lib as L
    D.tune_clear()
    D.tune_load(_cache(tmp_path, shape, 1, variant, tc))
    try:
        a = O.fill(23, shape, np.complex128)
        A = D.DeviceArray.from_host(a)
        p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
        got = {q["axis"]: q["variant"] for q in p.passes()}
        F = p.execute(A, A.similar())
        e = O.rel_l2(F.to_host(), O.fftn(a))
        print(shape, "variant", variant, "tc", tc, "passes", got, "rel-L2", e)
        assert e <= 1e-12
        pi = D.Plan(shape, L.C2C, +1, L.F64, D.Pencil(), (1, 2, 3))
        B = pi.execute(F, F.similar())
        assert O.rel_l2(B.to_host(), a) <= 1e-12
        # f32 through the same variant
        D.tune_load(_cache(tmp_path, shape, 0, variant, tc))
        a32 = O.fill(24, shape, np.complex64)
        F32 = D.fft(D.DeviceArray.from_host(a32))
        assert O.rel_l2(F32.to_host().astype(np.complex128), O.fftn(a32.astype(np.complex128))) <= 1e-5
    finally:
        D.tune_clear()


def test_default_uses_fixed_schedule(gpu):
    import daggerfft as D
    from daggerfft import _lib as L
    D.tune_clear()
    p = D.Plan((512, 256, 100), L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    v = [q["variant"] for q in p.passes()]
    # 512 (contiguous, 8 values per thread), 256 and 100 (strided, split LDS)
    # all run catalog kernels
    assert v == [1, 2, 2], v
