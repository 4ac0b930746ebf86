"""Plans created, executed and destroyed from several host threads at once:
the C ABI's plan_create/destroy are thread-safe and dfft_execute is
re-entrant for distinct (workspace, stream) pairs.  Eight threads, each on
its own stream, run a mix of shapes (catalog, runtime, Bluestein, long
lines, R2C) several times; every result equals the oracle's
(ComplexF64 rel-L2 <= 1e-12)."""
import threading

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [((64, 48, 40), "c2c"), ((11, 13, 17), "c2c"), ((2048, 3, 2), "c2c"), ((30, 20, 10), "r2c"),
         ((22, 9, 8), "r2c"), ((1050, 4, 3), "c2c"), ((127, 2, 5), "c2c"), ((96, 96, 96), "c2c")]


def test_concurrent_plans_and_streams(gpu):
    import torch
    import daggerfft as D
    from daggerfft import _lib as L
    refs = {}
    for k, (shape, kind) in enumerate(CASES):
        if kind == "c2c":
            a = O.fill(40 + k, shape, np.complex128)
            refs[k] = (a, O.fftn(a))
        else:
            a = O.fill(40 + k, shape, np.float64)
            refs[k] = (a, O.rfftn(a))
    errors = []

    def worker(t):
        try:
            s = torch.cuda.Stream()
            for rep in range(3):
                for k in range(len(CASES)):
                    shape, kind = CASES[(k + t) % len(CASES)]
                    a, ref = refs[(k + t) % len(CASES)]
                    with torch.cuda.stream(s):
                        X = D.DeviceArray.from_host(a)
                        if kind == "c2c":
                            p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
                            Y = D.DeviceArray(shape, np.complex128)
                        else:
                            p = D.Plan(shape, L.R2C, -1, L.F64, D.Pencil(), (1, 2, 3))
                            Y = D.DeviceArray((shape[0] // 2 + 1,) + shape[1:], np.complex128)
                        p.execute(X, Y, stream=s)
                    s.synchronize()
                    e = O.rel_l2(Y.to_host(), ref)
                    p.destroy()
                    if not e <= 1e-12:
                        errors.append((t, rep, shape, kind, e))
        except Exception as ex:  # reported below, in the test's own thread
            errors.append((t, repr(ex)))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(120)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors[:5]


def test_irfft_cached_plan_on_two_streams(gpu):
    """irfft through the process-wide plan cache from two streams at once:
    each call gets its own C2R workspace, so the two results equal the
    oracle's (the plan's own workspace would be shared by both)."""
    import torch
    import daggerfft as D
    shape = (96, 80, 64)
    refs, specs = [], []
    for k in range(2):
        a = O.fill(60 + k, shape, np.float64)
        spec = O.rfftn(a)
        refs.append(O.irfftn(spec, shape[0]))
        specs.append(D.DeviceArray.from_host(spec))
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for rep in range(5):
        outs = []
        for k in range(2):
            outs.append(D.irfft(specs[k], shape[0], stream=streams[k]))
        for s in streams:
            s.synchronize()
        for k in range(2):
            e = O.rel_l2(outs[k].to_host(), refs[k])
            assert e <= 1e-12, (rep, k, e)
