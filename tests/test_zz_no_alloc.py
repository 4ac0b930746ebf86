"""dfft_execute_static, the allocation-free variant of dfft_execute: through
the C ABI it must give bit-identical results to dfft_execute of the same plan
on the same input, on consecutive calls that share the device's static
scratch buffer (the second call smaller than the first, so it runs over what
the first left there), for plans with a workspace (C2R) and without (C2C,
R2C), and for a Slab plane-pipeline C2R plan.  A plan whose workspace exceeds
the scratch buffer is refused before anything is launched."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _static(plan, X, Y):
    import daggerfft as D
    from daggerfft import _lib as L
    st = L.lib().dfft_execute_static(plan.handle, ctypes.c_void_p(X.ptr), ctypes.c_void_p(Y.ptr),
                                     D._stream_handle(None))
    L.check(st, "dfft_execute_static")
    return Y


def _c2r_case(shape, prec, seed, flags=0, decomp=None):
    """A C2R plan, its half-spectrum input (the rfft of seeded data) and the
    result of dfft_execute with a caller workspace."""
    import torch
    import daggerfft as D
    from daggerfft import _lib as L
    decomp = decomp or D.Pencil()
    rdt = np.float64 if prec == L.F64 else np.float32
    x = D.rand(shape, rdt, seed=seed)
    F = D.rfft(x)
    p = D.Plan(shape, L.C2R, +1, prec, decomp, (1, 2, 3), flags=flags)
    assert p.workspace_bytes > 0
    ref = D.DeviceArray(shape, rdt)
    work = torch.empty(p.workspace_bytes, dtype=torch.uint8, device="cuda")
    p.execute(F, ref, work=work)
    torch.cuda.synchronize()
    return p, F, ref, x


def test_static_c2r_two_calls_second_smaller(gpu):
    import torch
    import daggerfft as D
    from daggerfft import _lib as L
    # call 1: 256^3 Float64 C2R (workspace 129*256*256*16 B = 135 MB of the 256 MiB)
    p1, F1, ref1, x1 = _c2r_case((256, 256, 256), L.F64, 2)
    assert p1.workspace_bytes <= L.STATIC_SCRATCH_BYTES
    y1 = _static(p1, F1, D.DeviceArray(p1.shape, np.float64))
    torch.cuda.synchronize()
    assert torch.equal(y1.tensor, ref1.tensor)
    e = float(torch.linalg.vector_norm(y1.tensor - x1.tensor) / torch.linalg.vector_norm(x1.tensor))
    print("256^3 F64 C2R static: bit-identical to dfft_execute, round trip rel-L2 %.2e" % e)
    assert e <= 1e-12
    # call 2: smaller, different shape and precision, over the first call's leftovers
    p2, F2, ref2, x2 = _c2r_case((60, 48, 30), L.F32, 5)
    assert p2.workspace_bytes < p1.workspace_bytes
    y2 = _static(p2, F2, D.DeviceArray(p2.shape, np.float32))
    torch.cuda.synchronize()
    assert torch.equal(y2.tensor, ref2.tensor)
    # and the oracle on the small one
    a = O.fill(5, (60, 48, 30), np.float32).astype(np.float64)
    e2 = O.rel_l2(y2.to_host().astype(np.float64), a)
    print("60x48x30 F32 C2R static: bit-identical, round trip rel-L2 %.2e" % e2)
    assert e2 <= 1e-5
    # call 3: the first plan again
    y3 = _static(p1, F1, D.DeviceArray(p1.shape, np.float64))
    torch.cuda.synchronize()
    assert torch.equal(y3.tensor, ref1.tensor)


def test_static_without_workspace_and_plane_pipeline(gpu):
    import torch
    import daggerfft as D
    from daggerfft import _lib as L
    # C2C and R2C use no workspace: the same launches as dfft_execute
    shape = (96, 80, 64)
    X = D.rand(shape, np.complex128, seed=7)
    p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    assert p.workspace_bytes == 0
    ref = p.execute(X, X.similar())
    got = _static(p, X, X.similar())
    R = D.rand(shape, np.float32, seed=8)
    pr = D.Plan(shape, L.R2C, -1, L.F32, D.Pencil(), (1, 2, 3))
    refr = pr.execute(R, D.DeviceArray((49, 80, 64), np.complex64))
    gotr = _static(pr, R, D.DeviceArray((49, 80, 64), np.complex64))
    torch.cuda.synchronize()
    assert torch.equal(got.tensor, ref.tensor) and torch.equal(gotr.tensor, refr.tensor)
    # a Slab plane-pipeline C2R plan: its grouped launches read the workspace too
    pp, Fp, refp, _ = _c2r_case((64, 48, 40), L.F64, 9, flags=L.FLAG_PLANE_PIPELINE, decomp=D.Slab())
    assert pp.schedule()[0] == "plane_pipeline"
    yp = _static(pp, Fp, D.DeviceArray(pp.shape, np.float64))
    torch.cuda.synchronize()
    assert torch.equal(yp.tensor, refp.tensor)
    # a graph plan runs its passes as plain launches here
    pg = D.Plan((40, 30, 20), L.C2R, +1, L.F64, D.Pencil(), (1, 2, 3), flags=L.FLAG_GRAPH)
    xg = D.rand((40, 30, 20), np.float64, seed=10)
    Fg = D.rfft(xg)
    yg = _static(pg, Fg, D.DeviceArray((40, 30, 20), np.float64))
    torch.cuda.synchronize()
    e = float(torch.linalg.vector_norm(yg.tensor - xg.tensor) / torch.linalg.vector_norm(xg.tensor))
    assert e <= 1e-12


def test_static_refuses_oversized_workspace(gpu):
    import daggerfft as D
    from daggerfft import _lib as L
    # 512x256x256 Float64 C2R: workspace 257*256*256*16 B > 256 MiB; the
    # call fails before any launch (full-size arrays all the same)
    p = D.Plan((512, 256, 256), L.C2R, +1, L.F64, D.Pencil(), (1, 2, 3))
    assert p.workspace_bytes > L.STATIC_SCRATCH_BYTES
    X = D.DeviceArray((257, 256, 256), np.complex128)
    Y = D.DeviceArray((512, 256, 256), np.float64)
    st = L.lib().dfft_execute_static(p.handle, ctypes.c_void_p(X.ptr), ctypes.c_void_p(Y.ptr),
                                     D._stream_handle(None))
    assert st == L.DFFT_ERR_SCRATCH
    assert b"exceeds the static scratch" in L.lib().dfft_last_error()
    assert L.lib().dfft_execute_static(None, None, None, None) == L.DFFT_ERR_ARG
