"""The serving configuration at its real sizes: the default trace (2,000
requests, Poisson arrivals, 60/40 C2C/R2C, cubic edges 64..512 log-uniform,
ComplexF32/ComplexF64 50/50, seed 8) replayed saturating through one server
with a 200 GiB HBM budget, 4 lanes and 32 requests in flight.

Kept results: ten seeded requests from every decile of the edge
distribution, plus every 512^3 ComplexF64 C2C request and the largest R2C
requests of each precision (edges >= 448).  Each kept result must be
bit-identical to the same transform run alone through the plain API (same
plan, same kernels); results up to 128^3 are compared in full with the oracle
(ComplexF64 rel-L2 <= 1e-12, ComplexF32 against the FP64 oracle <= 1e-5) and
the three largest against the oracle's long-double spot bins."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _keep(trace):
    from daggerfft import _lib as L
    order = sorted(range(len(trace)), key=lambda r: (trace[r].shape[0], r))
    rng = np.random.default_rng(8)
    keep = set()
    for d in range(10):
        dec = order[d * len(order) // 10:(d + 1) * len(order) // 10]
        keep.update(int(r) for r in rng.choice(dec, size=10, replace=False))
    keep.update(r for r, q in enumerate(trace) if q.shape[0] == 512 and q.kind == L.C2C and q.precision == L.F64)
    for prec in (L.F32, L.F64):
        r2c = [r for r, q in enumerate(trace) if q.kind == L.R2C and q.precision == prec and q.shape[0] >= 448]
        keep.update(sorted(r2c, key=lambda r: -trace[r].shape[0])[:2])
    return sorted(keep)


def _spot(q, y, nbins=10):
    from daggerfft import _lib as L
    n = q.shape
    rng = np.random.default_rng(q.seed)
    n0 = y.shape[0]
    bins = [(0, 0, 0), (n0 - 1, n[1] // 2, n[2] - 1)] + [
        (int(rng.integers(n0)), int(rng.integers(n[1])), int(rng.integers(n[2]))) for _ in range(nbins)]
    x = O.fill(q.seed, n, np.dtype(q.in_dtype)).astype(np.complex128)
    ref = O.spot_bins(x, bins)
    got = np.array([y[b] for b in bins]).astype(np.complex128)
    return O.rel_l2(got, ref)


def test_serving_trace_full_size(gpu):
    import torch
    import daggerfft as D
    from daggerfft import _lib as L
    from daggerfft import serve as S
    trace = S.serving_trace()
    assert len(trace) == 2000
    keep = _keep(trace)
    edges = sorted(trace[r].shape[0] for r in keep)
    print("keeping %d results, edges %s" % (len(keep), edges), flush=True)
    assert len(keep) >= 100
    assert any(trace[r].shape[0] == 512 and trace[r].kind == L.C2C and trace[r].precision == L.F64 for r in keep)
    assert any(trace[r].kind == L.R2C and trace[r].shape[0] >= 448 for r in keep)
    with S.Server(hbm_budget=200 * S.GiB, max_inflight=32, lanes=4) as srv:
        res = S.replay(srv, trace, open_loop=False, keep=keep)
        st = srv.stats()
    print("saturating: %.0f req/s, %.0f GB/s, peak running %d, failed %d"
          % (res["req_per_s"], res["GBps"], st["peak_running"], st["failed"]), flush=True)
    assert st["completed"] == 2000 and st["failed"] == 0
    assert st["peak_running"] >= 2
    got = res["results"]
    assert sorted(got) == keep
    torch.cuda.synchronize()
    # bit-identical to the same transform alone
    for r in keep:
        q = trace[r]
        x = D.rand(q.shape, q.in_dtype, seed=q.seed)
        y = D.rfft(x) if q.kind == L.R2C else D.fft(x)
        torch.cuda.synchronize()
        assert torch.equal(y.tensor, got[r].tensor), (r, q)
        del x, y
    print("all %d kept results bit-identical to the plain call alone" % len(keep), flush=True)
    # the oracle: every element of the results up to 128^3
    small = [r for r in keep if trace[r].shape[0] <= 128]
    assert len(small) >= 20
    worst = {L.F32: 0.0, L.F64: 0.0}
    for r in small:
        q = trace[r]
        x = O.fill(q.seed, q.shape, np.dtype(q.in_dtype))
        ref = O.rfftn(x.astype(np.float64)) if q.kind == L.R2C else O.fftn(x.astype(np.complex128))
        e = O.rel_l2(got[r].to_host().astype(np.complex128), ref)
        worst[q.precision] = max(worst[q.precision], e)
        assert e <= (1e-12 if q.precision == L.F64 else 1e-5), (r, q, e)
    print("oracle, %d results <= 128^3: worst rel-L2 F64 %.2e, F32 %.2e" % (len(small), worst[L.F64], worst[L.F32]))
    # spot bins of the three largest
    for r in sorted(keep, key=lambda r: (-trace[r].shape[0], r))[:3]:
        q = trace[r]
        e = _spot(q, got[r].to_host())
        print("spot bins of request %d (%d^3, kind %d, prec %d): rel-L2 %.2e" % (r, q.shape[0], q.kind, q.precision, e))
        assert e <= (1e-12 if q.precision == L.F64 else 1e-5), (r, q, e)
