"""A batch as ONE cooperative launch (DFFT_FLAG_BATCH_PERSISTENT): the
grouped schedule's tiles, walked by as many workgroups as the device holds at
once, with a grid barrier between axis-pass levels.  It runs the same tile
code on the same schedules as the three grouped launches, so its results
equal theirs bit for bit; the oracle checks a subset (ComplexF64 rel-L2 <=
1e-12, ComplexF32 vs the FP64 oracle <= 1e-5).

The file is named to be collected after the other GPU tests (and before the
no-allocation entry point's): after a process's first cooperative launch,
work spread over several streams runs slower for the rest of the process,
which would slow the multi-stream tests that follow in it."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _inputs(descs, seeds):
    import daggerfft as D
    ins = []
    for (shape, kind, sign, prec), sd in zip(descs, seeds):
        if kind == 0:
            ins.append(D.rand(shape, np.complex128 if prec == 1 else np.complex64, seed=sd))
        else:
            ins.append(D.rand(shape, np.float64 if prec == 1 else np.float32, seed=sd))
    return ins


def _run(b, descs, ins):
    import daggerfft as D
    outs = []
    for shape, kind, sign, prec in descs:
        dt = np.complex128 if prec == 1 else np.complex64
        oshape = shape if kind == 0 else (shape[0] // 2 + 1,) + tuple(shape[1:])
        outs.append(D.DeviceArray(oshape, dt))
    b.execute(ins, outs)
    return [o.to_host() for o in outs]


def _small_mix():
    from daggerfft import _lib as L
    return [((32, 48, 20), L.C2C, +1, 1), ((64, 30, 14), L.R2C, -1, 1), ((16, 16, 16), L.C2C, -1, 0),
            ((40, 21, 10), L.R2C, -1, 0), ((7, 5, 3), L.C2C, -1, 1)]


def test_persistent_batch_small_mix_equals_grouped(gpu):
    """C2C both signs, R2C, both precisions, a runtime-engine transform
    (7x5x3): one launch, the grouped launches' results bit for bit, on three
    executes in a row (the grid barrier's state carries over launches)."""
    from daggerfft import _lib as L
    from daggerfft.batch import Batch
    descs = _small_mix()
    seeds = [51, 52, 53, 54, 55]
    print("persistent small mix: create", flush=True)
    p = Batch(descs, flags=L.FLAG_BATCH_PERSISTENT)
    g = Batch(descs, flags=L.FLAG_BATCH_GROUPED)
    assert p.launches == 1 and g.launches == 3
    assert p.tickets == g.tickets
    ins = _inputs(descs, seeds)
    yg = _run(g, descs, ins)
    for rep in range(3):
        print("persistent small mix: execute %d" % rep, flush=True)
        yp = _run(p, descs, ins)
        for t, (a, b) in enumerate(zip(yp, yg)):
            assert np.array_equal(a, b), (rep, t)


def test_persistent_batch_config_64_mixed(gpu):
    """The 64-transform mix (axis lengths from {32..256} by seed 7, half
    ComplexF32): one launch, equal to the grouped launches bit for bit, then a
    second, smaller batch (the first 16 transforms) through its own
    persistent launch; every 8th transform against the oracle."""
    from daggerfft import _lib as L
    from daggerfft.batch import Batch, batch_config_sizes
    cfg = batch_config_sizes(64, 7)
    descs = [(shape, L.C2C, -1, prec) for shape, prec in cfg]
    seeds = [1000 + t for t in range(64)]
    print("persistent 64 mix: create", flush=True)
    p = Batch(descs, flags=L.FLAG_BATCH_PERSISTENT)
    g = Batch(descs, flags=L.FLAG_BATCH_GROUPED)
    assert p.launches == 1
    ins = _inputs(descs, seeds)
    yg = _run(g, descs, ins)
    print("persistent 64 mix: execute", flush=True)
    yp = _run(p, descs, ins)
    worst64 = worst32 = 0.0
    for t, ((shape, prec), a, b) in enumerate(zip(cfg, yp, yg)):
        assert np.array_equal(a, b), t
        if t % 8 == 0:
            dt = np.complex128 if prec == 1 else np.complex64
            x = O.fill(1000 + t, shape, dt).astype(np.complex128)
            e = O.rel_l2(a.astype(np.complex128), O.fftn(x))
            if prec == 1:
                worst64 = max(worst64, e)
            else:
                worst32 = max(worst32, e)
    print("persistent batch of 64: worst rel-L2 f64 %.3e, f32 %.3e" % (worst64, worst32))
    assert worst64 <= 1e-12 and worst32 <= 1e-5
    # a smaller batch next: fewer tiles per level than the grid holds
    small = descs[:16]
    print("persistent 16 of the mix: execute", flush=True)
    p16 = Batch(small, flags=L.FLAG_BATCH_PERSISTENT)
    y16 = _run(p16, small, ins[:16])
    for t in range(16):
        assert np.array_equal(y16[t], yg[t]), t


def test_persistent_batch_flag_errors(gpu):
    """Creation refuses what one 256-thread cooperative launch cannot run:
    a level needing wider tiles (a 4096-point ComplexF32 contiguous axis) is
    DFFT_ERR_UNSUPPORTED_SIZE; with the lanes flag or a graph, DFFT_ERR_ARG.
    Nothing is launched."""
    from daggerfft import _lib as L
    from daggerfft.batch import Batch
    wide = [((4096, 4, 2), L.C2C, -1, 0), ((64, 64, 48), L.C2C, -1, 1)]
    with pytest.raises(L.DFFTError) as ei:
        Batch(wide, flags=L.FLAG_BATCH_PERSISTENT)
    assert ei.value.status == L.DFFT_ERR_UNSUPPORTED_SIZE
    for extra in (L.FLAG_BATCH_LANES, L.FLAG_GRAPH):
        with pytest.raises(L.DFFTError) as ei:
            Batch(_small_mix(), flags=L.FLAG_BATCH_PERSISTENT | extra)
        assert ei.value.status == L.DFFT_ERR_ARG


def test_persistent_batch_edge_mixes(gpu):
    """The single cooperative launch on a one-transform batch and on axes of
    length 1: equal to the grouped launches bit for bit."""
    from daggerfft import _lib as L
    from daggerfft.batch import Batch
    for descs in ([((64, 48, 32), L.C2C, -1, 1)],
                  [((1, 16, 1), L.C2C, -1, 1), ((2, 1, 3), L.C2C, +1, 0), ((32, 1, 48), L.C2C, -1, 0)]):
        print("persistent edge mix", [d[0] for d in descs], flush=True)
        p = Batch(descs, flags=L.FLAG_BATCH_PERSISTENT)
        g = Batch(descs, flags=L.FLAG_BATCH_GROUPED)
        assert p.launches == 1
        ins = _inputs(descs, [90 + t for t in range(len(descs))])
        for a, b in zip(_run(p, descs, ins), _run(g, descs, ins)):
            assert np.array_equal(a, b)
