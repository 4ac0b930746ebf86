"""Batched launch of 64 mixed-size 3D transforms: every result equals the
same transform planned and run alone up to rounding (the grouped kernel and
the plain path's catalog kernels order the same arithmetic differently:
ComplexF64 rel-L2 <= 1e-14, ComplexF32 <= 1e-6), and the oracle within
tolerance (ComplexF64 rel-L2 <= 1e-12; ComplexF32 vs the FP64 oracle <= 1e-5).

The first three tests exercised an earlier single persistent launch whose
workgroups waited on each other through device counters.  That kernel hung
on the GPU twice (the second time on the run meant to test a fix); no cause
was found, so the kernel was removed and these tests stay skipped.  The
batch now runs as grouped launches (one per axis-pass level, no waits
between workgroups), tested by the test_grouped_* functions below, or as
those launches' tiles in one cooperative launch with grid barriers between
levels (test_zy_persistent_batch.py)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _run_batch(descs, seeds):
    import daggerfft as D
    from daggerfft.batch import Batch
    b = Batch(descs)
    ins, outs = [], []
    for (shape, kind, sign, prec), sd in zip(descs, seeds):
        if kind == 0:
            dt = np.complex128 if prec == 1 else np.complex64
            ins.append(D.rand(shape, dt, seed=sd))
            outs.append(D.DeviceArray(shape, dt))
        else:
            dt = np.float64 if prec == 1 else np.float32
            ins.append(D.rand(shape, dt, seed=sd))
            oshape = (shape[0] // 2 + 1,) + tuple(shape[1:])
            outs.append(D.DeviceArray(oshape, np.complex128 if prec == 1 else np.complex64))
    b.execute(ins, outs)
    assert b.error_flag() == 0
    return b, ins, outs


@pytest.mark.skip(reason="persistent inter-workgroup-wait batch kernel hung on the GPU twice; removed")
def test_batch_config_64_mixed(gpu):
    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft.batch import batch_config_sizes
    cfg = batch_config_sizes(64, 7)
    descs = [(shape, L.C2C, -1, prec) for shape, prec in cfg]
    seeds = [1000 + t for t in range(64)]
    b, ins, outs = _run_batch(descs, seeds)
    worst64 = worst32 = 0.0
    for t, ((shape, prec), X, Y) in enumerate(zip(cfg, ins, outs)):
        alone = D.fft(X)
        yb = Y.to_host()
        assert np.array_equal(yb, alone.to_host()) or O.rel_l2(yb, alone.to_host()) < 1e-15, t
        if t % 8 == 0 or t < 4:  # oracle on a subset keeps the CPU time low
            dt = np.complex128 if prec == 1 else np.complex64
            a = O.fill(1000 + t, shape, dt).astype(np.complex128)
            e = O.rel_l2(yb.astype(np.complex128), O.fftn(a))
            if prec == 1:
                worst64 = max(worst64, e)
            else:
                worst32 = max(worst32, e)
    print("batch of 64: worst rel-L2 f64 %.3e, f32 %.3e (vs FP64 oracle), tickets %d" % (worst64, worst32, b.tickets))
    assert worst64 <= 1e-12 and worst32 <= 1e-5


@pytest.mark.skip(reason="persistent inter-workgroup-wait batch kernel hung on the GPU twice; removed")
def test_batch_inverse_and_r2c_mix(gpu):
    import daggerfft as D
    from daggerfft import _lib as L
    descs = [((32, 48, 20), L.C2C, +1, 1), ((64, 30, 14), L.R2C, -1, 1), ((16, 16, 16), L.C2C, -1, 0),
             ((40, 21, 10), L.R2C, -1, 0), ((7, 5, 3), L.C2C, -1, 1)]
    seeds = [51, 52, 53, 54, 55]
    b, ins, outs = _run_batch(descs, seeds)
    for (shape, kind, sign, prec), X, Y in zip(descs, ins, outs):
        if kind == L.C2C:
            alone = D.fft(X) if sign < 0 else D.ifft(X)
        else:
            alone = D.rfft(X)
        tol = 1e-14 if prec == 1 else 1e-6  # grouped vs catalog kernels: rounding only
        assert O.rel_l2(Y.to_host(), alone.to_host()) <= tol


@pytest.mark.skip(reason="persistent inter-workgroup-wait batch kernel hung on the GPU twice; removed")
def test_batch_spin_limit_reports_error(gpu):
    """A tiny spin limit forced from the host makes dependency waits give up:
    the kernel still drains and the error flag is raised."""
    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft.batch import Batch
    descs = [((256, 256, 64), L.C2C, -1, 1)] * 2
    b = Batch(descs)
    b.set_spin_limit(1)
    ins = [D.rand(d[0], np.complex128, seed=3) for d in descs]
    outs = [D.DeviceArray(d[0], np.complex128) for d in descs]
    b.execute(ins, outs)
    assert b.error_flag() == 1
    b.set_spin_limit(1 << 32)
    b.execute(ins, outs)
    assert b.error_flag() == 0
    assert O.rel_l2(outs[0].to_host(), D.fft(ins[0]).to_host()) < 1e-15


def _run_grouped(descs, seeds):
    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft.batch import Batch
    b = Batch(descs, flags=L.FLAG_BATCH_GROUPED)
    # one launch per level, two where a level of compile-time schedules holds both precisions
    assert 3 <= b.launches <= 6
    ins, outs = [], []
    for (shape, kind, sign, prec), sd in zip(descs, seeds):
        if kind == 0:
            dt = np.complex128 if prec == 1 else np.complex64
            ins.append(D.rand(shape, dt, seed=sd))
            outs.append(D.DeviceArray(shape, dt))
        else:
            dt = np.float64 if prec == 1 else np.float32
            ins.append(D.rand(shape, dt, seed=sd))
            oshape = (shape[0] // 2 + 1,) + tuple(shape[1:])
            outs.append(D.DeviceArray(oshape, np.complex128 if prec == 1 else np.complex64))
    b.execute(ins, outs)
    return b, ins, outs


def test_grouped_batch_small_mix(gpu):
    import daggerfft as D
    from daggerfft import _lib as L
    descs = [((32, 48, 20), L.C2C, +1, 1), ((64, 30, 14), L.R2C, -1, 1), ((16, 16, 16), L.C2C, -1, 0),
             ((40, 21, 10), L.R2C, -1, 0), ((7, 5, 3), L.C2C, -1, 1)]
    b, ins, outs = _run_grouped(descs, [51, 52, 53, 54, 55])
    for (shape, kind, sign, prec), X, Y in zip(descs, ins, outs):
        if kind == L.C2C:
            alone = D.fft(X) if sign < 0 else D.ifft(X)
        else:
            alone = D.rfft(X)
        tol = 1e-14 if prec == 1 else 1e-6  # grouped vs catalog kernels: rounding only
        assert O.rel_l2(Y.to_host(), alone.to_host()) <= tol


def test_grouped_batch_wide_level(gpu):
    """A 4096-point ComplexF32 contiguous pass needs 512-thread tiles: its level runs the
    runtime engine throughout, while the other levels run the compile-time
    schedules (the 64- and 48-point passes) beside runtime ones (4, 2, 30)."""
    import daggerfft as D
    from daggerfft import _lib as L
    descs = [((4096, 4, 2), L.C2C, -1, 0), ((64, 64, 48), L.C2C, -1, 1), ((48, 30, 64), L.C2C, +1, 0)]
    b, ins, outs = _run_grouped(descs, [61, 62, 63])
    for t, ((shape, kind, sign, prec), X, Y) in enumerate(zip(descs, ins, outs)):
        dt = np.complex128 if prec == 1 else np.complex64
        a = O.fill(61 + t, shape, dt).astype(np.complex128)
        ref = O.fftn(a, inverse=sign > 0)
        e = O.rel_l2(Y.to_host().astype(np.complex128), ref)
        assert e <= (1e-12 if prec == 1 else 1e-5), (t, e)


def test_grouped_batch_config_64_mixed(gpu):
    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft.batch import batch_config_sizes
    cfg = batch_config_sizes(64, 7)
    descs = [(shape, L.C2C, -1, prec) for shape, prec in cfg]
    b, ins, outs = _run_grouped(descs, [1000 + t for t in range(64)])
    worst64 = worst32 = 0.0
    for t, ((shape, prec), X, Y) in enumerate(zip(cfg, ins, outs)):
        yb = Y.to_host()
        # the grouped kernel and the plain path's catalog kernel round differently
        assert O.rel_l2(yb, D.fft(X).to_host()) < (1e-14 if prec == 1 else 1e-6), t
        # every transform of the batch against the oracle
        dt = np.complex128 if prec == 1 else np.complex64
        a = O.fill(1000 + t, shape, dt).astype(np.complex128)
        e = O.rel_l2(yb.astype(np.complex128), O.fftn(a))
        assert e <= (1e-12 if prec == 1 else 1e-5), (t, shape, e)
        if prec == 1:
            worst64 = max(worst64, e)
        else:
            worst32 = max(worst32, e)
    print("grouped batch of 64: worst rel-L2 f64 %.3e, f32 %.3e, workgroups %d" % (worst64, worst32, b.tickets))
    assert worst64 <= 1e-12 and worst32 <= 1e-5


def test_grouped_batch_graph_replays(gpu):
    """With DFFT_FLAG_GRAPH the batch's launches are recorded once as a
    hipGraph and replayed: every replay -- new input data in the same
    buffers, the program's own eager work and synchronisations in between,
    then a second set of buffers (re-recorded) -- equals the eager launches
    bit for bit."""
    import ctypes

    import torch

    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft.batch import Batch
    descs = [((32, 48, 20), L.C2C, -1, 1), ((64, 30, 14), L.R2C, -1, 1), ((16, 16, 16), L.C2C, +1, 0),
             ((40, 21, 10), L.R2C, -1, 0), ((96, 32, 48), L.C2C, -1, 1)]
    eager = Batch(descs, flags=L.FLAG_BATCH_GROUPED)
    graph = Batch(descs, flags=L.FLAG_BATCH_GROUPED | L.FLAG_GRAPH)

    def buffers(seed0):
        ins, oe, og = [], [], []
        for t, (shape, kind, sign, prec) in enumerate(descs):
            cplx = kind == L.C2C
            dt = (np.complex128 if prec else np.complex64) if cplx else (np.float64 if prec else np.float32)
            ins.append(D.rand(shape, dt, seed=seed0 + t))
            oshape = shape if cplx else (shape[0] // 2 + 1,) + tuple(shape[1:])
            odt = np.complex128 if prec else np.complex64
            oe.append(D.DeviceArray(oshape, odt))
            og.append(D.DeviceArray(oshape, odt))
        return ins, oe, og

    def refill(ins, seed0):
        for t, X in enumerate(ins):
            L.check(L.lib().dfft_fill_uniform(ctypes.c_void_p(X.ptr), X.size, 1 if X.is_complex else 0,
                                              X.precision, ctypes.c_uint64(seed0 + t),
                                              D._stream_handle(None)))

    scratch = torch.rand(1 << 20, device="cuda")
    for bufset in range(2):
        ins, oe, og = buffers(70 + 10 * bufset)
        for rep in range(3):
            refill(ins, 100 * (bufset + 1) + rep)
            eager.execute(ins, oe)
            torch.cuda.synchronize()
            scratch.mul_(1.5).add_(0.25)  # eager work between replays
            graph.execute(ins, og)
            scratch.sum().item()
            torch.cuda.synchronize()
            for t, (a, b) in enumerate(zip(oe, og)):
                assert torch.equal(a.tensor, b.tensor), (bufset, rep, t)
    # and the replays are right, not merely repeatable
    for t, ((shape, kind, sign, prec), X, Y) in enumerate(zip(descs, ins, og)):
        x = X.to_host().astype(np.complex128 if kind == L.C2C else np.float64)
        ref = (O.fftn(x) if sign < 0 else O.fftn(x, inverse=True)) if kind == L.C2C else O.rfftn(x)
        assert O.rel_l2(Y.to_host().astype(np.complex128), ref) <= (1e-12 if prec else 1e-5), t


@pytest.mark.parametrize("flags", ["lanes", "lanes_graph"])
def test_lanes_batch_equals_plans_bitwise(gpu, flags):
    """The lanes schedule: each transform runs its own plan on one of four
    forked streams (optionally recorded as one hipGraph with parallel
    branches); results equal the plain plan of each transform bit for bit,
    over replays with new data and eager work in between."""
    import ctypes

    import torch

    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft.batch import Batch, batch_config_sizes
    cfg = batch_config_sizes(24, 9)
    descs = [(shape, L.C2C, -1 if t % 3 else +1, prec) for t, (shape, prec) in enumerate(cfg)]
    descs += [((64, 30, 14), L.R2C, -1, 1), ((40, 21, 10), L.R2C, -1, 0)]
    fl = L.FLAG_BATCH_LANES | (L.FLAG_GRAPH if flags == "lanes_graph" else 0)
    b = Batch(descs, flags=fl)
    assert b.launches == 3 * len(descs)
    ins, outs = [], []
    for t, (shape, kind, sign, prec) in enumerate(descs):
        cplx = kind == L.C2C
        dt = (np.complex128 if prec else np.complex64) if cplx else (np.float64 if prec else np.float32)
        ins.append(D.rand(shape, dt, seed=300 + t))
        oshape = shape if cplx else (shape[0] // 2 + 1,) + tuple(shape[1:])
        outs.append(D.DeviceArray(oshape, np.complex128 if prec else np.complex64))
    scratch = torch.rand(1 << 20, device="cuda")
    for rep in range(3):
        for t, X in enumerate(ins):
            L.check(L.lib().dfft_fill_uniform(ctypes.c_void_p(X.ptr), X.size, 1 if X.is_complex else 0,
                                              X.precision, ctypes.c_uint64(500 + 50 * rep + t),
                                              D._stream_handle(None)))
        b.execute(ins, outs)
        scratch.mul_(1.5).add_(0.25)
        torch.cuda.synchronize()
        for t, ((shape, kind, sign, prec), X, Y) in enumerate(zip(descs, ins, outs)):
            if kind == L.R2C:
                alone = D.rfft(X)
            else:
                alone = D.fft(X) if sign < 0 else D.ifft(X)
            assert torch.equal(Y.tensor, alone.tensor), (rep, t)


def test_batch_default_schedule(gpu):
    """Without a schedule flag a batch takes the schedule cache's measured
    pick for its mix (the shipped cache picks the grouped launches for the
    64-transform mix: six, one per level and precision); with no pick it runs grouped when every transform is C2C and every
    axis length has a compile-time grouped schedule (one launch per level and
    precision: six for the mix of both precisions), else lanes (three
    launches per transform); both flags at once are refused."""
    import ctypes
    import os

    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft.batch import Batch, batch_config_sizes
    D.tune_load(os.path.join(os.path.dirname(D.__file__), "..", "wisdom", "gfx950.txt"))
    cfg = batch_config_sizes(64, 7)
    mix = [(shape, L.C2C, -1, prec) for shape, prec in cfg]
    assert Batch(mix).launches == 6
    assert Batch(mix, flags=L.FLAG_BATCH_LANES).launches == 192
    assert Batch(mix, flags=L.FLAG_NO_TUNE).launches == 6
    assert Batch([((64, 30, 48), L.C2C, -1, 1), ((32, 32, 32), L.C2C, -1, 0)]).launches == 6
    assert Batch([((64, 64, 64), L.R2C, -1, 1), ((32, 32, 32), L.C2C, -1, 0)]).launches == 6
    arr = (L.dfft_batch_desc * 1)()
    arr[0].n[0], arr[0].n[1], arr[0].n[2] = 32, 32, 32
    arr[0].kind, arr[0].sign, arr[0].precision = L.C2C, -1, 1
    h = ctypes.c_void_p()
    st = L.lib().dfft_batch_create(ctypes.byref(h), 1, arr, L.FLAG_BATCH_GROUPED | L.FLAG_BATCH_LANES)
    assert st == L.DFFT_ERR_ARG


def test_measured_batch_schedule_is_cached(gpu, tmp_path):
    """DFFT_FLAG_MEASURE on a batch with no schedule flag times the grouped
    and lanes schedules and caches the faster under the mix's key: a second
    batch of the same mix without the flag takes the cached schedule, the
    saved cache holds the entry, and the results equal the oracle's."""
    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft.batch import Batch
    descs = [((64, 48, 32), L.C2C, -1, 1), ((32, 32, 96), L.C2C, -1, 0), ((96, 64, 48), L.C2C, +1, 1),
             ((48, 48, 48), L.C2C, -1, 0), ((128, 32, 64), L.C2C, -1, 1)]
    import os
    shipped = os.path.join(os.path.dirname(D.__file__), "..", "wisdom", "gfx950.txt")
    D.tune_clear()
    D.tune_load(shipped)
    try:
        n0 = D.tune_entries()
        b = Batch(descs, flags=L.FLAG_MEASURE)
        picked = b.launches
        assert picked in (6, 15), picked  # grouped: 3 levels x 2 precisions; lanes: 3 per transform
        assert D.tune_entries() > n0
        again = Batch(descs)
        assert again.launches == picked
        path = str(tmp_path / "wisdom.txt")
        D.tune_save(path)
        assert any(line.startswith("batch/5/") for line in open(path))
        ins = [D.rand(shape, np.complex128 if prec == 1 else np.complex64, seed=70 + t)
               for t, (shape, _, _, prec) in enumerate(descs)]
        outs = [X.similar() for X in ins]
        again.execute(ins, outs)
        for t, ((shape, kind, sign, prec), Y) in enumerate(zip(descs, outs)):
            dt = np.complex128 if prec == 1 else np.complex64
            a = O.fill(70 + t, shape, dt).astype(np.complex128)
            e = O.rel_l2(Y.to_host().astype(np.complex128), O.fftn(a, inverse=sign > 0))
            assert e <= (1e-12 if prec == 1 else 1e-5), (t, e)
    finally:
        D.tune_clear()
        D.tune_load(shipped)


@pytest.mark.parametrize("descs", [
    [((64, 48, 32), 0, -1, 1)],                                      # one transform
    [((1, 16, 1), 0, -1, 1), ((2, 1, 3), 0, +1, 0), ((32, 1, 48), 0, -1, 0)],  # axes of length 1
    [((96, 64, 32), 0, -1, 0), ((96, 64, 32), 0, -1, 1), ((30, 64, 32), 0, -1, 1)],  # levels split and not
])
def test_grouped_batch_edge_mixes(gpu, descs):
    """Grouped launches (one per level, or per level and precision where the
    level's passes all run compile-time schedules) on a single transform,
    axes of length 1, and a mix whose first level holds a runtime-engine pass
    (30 points) while the others split by precision: every result within
    rounding of the same transform planned alone, and the oracle."""
    import daggerfft as D
    b, ins, outs = _run_grouped(descs, [90 + t for t in range(len(descs))])
    for t, ((shape, kind, sign, prec), X, Y) in enumerate(zip(descs, ins, outs)):
        alone = D.fft(X) if sign < 0 else D.ifft(X)
        assert O.rel_l2(Y.to_host(), alone.to_host()) <= (1e-14 if prec == 1 else 1e-6), t
        dt = np.complex128 if prec == 1 else np.complex64
        a = O.fill(90 + t, shape, dt).astype(np.complex128)
        e = O.rel_l2(Y.to_host().astype(np.complex128), O.fftn(a, inverse=sign > 0))
        assert e <= (1e-12 if prec == 1 else 1e-5), (t, shape, e)
