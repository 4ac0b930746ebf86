"""Wisdom on the device: DFFT_FLAG_MEASURE tunes a plan's untuned passes at
creation and the cache is saved to the autosave path (FFTW.MEASURE plus
save_wisdom() after every new plan, src/fft.jl:127,162); a plan's exported
wisdom, loaded into an empty cache, rebuilds the identical plan (same
kernels, radix lists, tile shapes and launch geometry) whose results equal
the oracle (ComplexF64 rel-L2 <= 1e-12)."""
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as O

pytestmark = pytest.mark.gpu
SHIPPED = os.path.join(ROOT, "daggerfft.jl_amd", "wisdom", "gfx950.txt")
KEYS = ("kernel", "radices", "TPL", "TC", "threads", "blocks", "lds", "swz", "variant", "bluestein")


def _restore(D):
    D.tune_autosave(None)
    D.tune_clear()
    D.tune_load(SHIPPED)


def test_measure_flag_tunes_and_autosaves(gpu, tmp_path):
    import daggerfft as D
    from daggerfft import _lib as L
    path = str(tmp_path / "auto.txt")
    shape = (96, 80, 60)
    try:
        D.tune_clear()
        D.tune_autosave(path)
        p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3), flags=L.FLAG_MEASURE)
        assert D.tune_entries() == 3  # one entry per distinct axis pass
        lines = open(path).read().splitlines()
        assert lines[0].startswith("#") and "hip=" in lines[0] and "arch=gfx950" in lines[0]
        assert len([ln for ln in lines if not ln.startswith("#")]) == 3
        a = O.fill(41, shape, np.complex128)
        F = p.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, np.complex128))
        assert O.rel_l2(F.to_host(), O.fftn(a)) <= 1e-12
        # a second MEASURE plan of the same shape finds every pass tuned
        D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3), flags=L.FLAG_MEASURE)
        assert D.tune_entries() == 3
    finally:
        _restore(D)


@pytest.mark.parametrize("shape,kind", [((512, 512, 64), "c2c"), ((384, 240, 210), "c2c"), ((127, 64, 30), "c2c"),
                                        ((256, 128, 96), "r2c"), ((1050, 24, 20), "c2c")])
def test_plan_wisdom_rebuilds_identical_plan(gpu, tmp_path, shape, kind):
    import daggerfft as D
    from daggerfft import _lib as L
    path = str(tmp_path / "plan.txt")
    k = L.C2C if kind == "c2c" else L.R2C
    try:
        D.tune_clear()
        D.tune_load(SHIPPED)
        p1 = D.Plan(shape, k, -1, L.F64, D.Slab(), (1, 2, 3))
        d1 = [{q: x[q] for q in KEYS} for x in p1.passes()]
        p1.export_wisdom(path)
        D.tune_clear()
        D.tune_load(path)
        p2 = D.Plan(shape, k, -1, L.F64, D.Slab(), (1, 2, 3))
        d2 = [{q: x[q] for q in KEYS} for x in p2.passes()]
        assert d1 == d2
        if kind == "c2c":
            a = O.fill(42, shape, np.complex128)
            F = p2.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, np.complex128))
            assert O.rel_l2(F.to_host(), O.fftn(a)) <= 1e-12
        else:
            a = O.fill(42, shape, np.float64)
            F = p2.execute(D.DeviceArray.from_host(a), D.DeviceArray((shape[0] // 2 + 1,) + shape[1:], np.complex128))
            assert O.rel_l2(F.to_host(), O.rfftn(a)) <= 1e-12
    finally:
        _restore(D)
