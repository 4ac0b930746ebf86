"""Serving loop (dfft_server_*): every request's result is the same alone and
in the mix, admission respects the HBM budget, and caller buffers, cancels
and errors behave as documented.

"Alone" is the plain API (daggerfft.fft / rfft of daggerfft.rand(seed)) run on
its own with nothing else on the device; "in the mix" is the same request
served while other requests of other sizes, kinds and precisions are queued
and running on other streams.  Both run the same plan of the same shape, so
the results are expected to be bit-identical; the test's bar is the stated
accuracy (ComplexF64 rel-L2 <= 1e-12, ComplexF32 <= 1e-6 between the two),
and a few requests are also checked against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {0: 1e-6, 1: 1e-12}  # precision -> rel-L2 bound between served and alone


def _alone(q):
    import daggerfft as D
    from daggerfft import _lib as L
    x = D.rand(q.shape, q.in_dtype, seed=q.seed)
    y = D.rfft(x) if q.kind == L.R2C else D.fft(x)
    return y.to_host()


def _mix(seed=3, count=24, lo=16, hi=96):
    from daggerfft import serve as S
    return S.serving_trace(count=count, seed=seed, rate=4000.0, lo=lo, hi=hi)


def test_served_mix_equals_alone(gpu):
    from daggerfft import serve as S
    trace = _mix()
    print("mix:", [(q.shape[0], q.kind, q.precision) for q in trace], flush=True)
    with S.Server(hbm_budget=4 * S.GiB, max_inflight=8, lanes=4) as srv:
        tickets = [srv.submit(q) for q in trace]
        got = []
        for t, q in zip(tickets, trace):
            assert srv.wait(t, 60.0)
            inf = srv.info(t)
            assert inf["status"] == 0, inf
            assert inf["t_done"] >= inf["t_admit"] >= inf["t_arrival"] - 1e-9
            got.append(srv.result(t, q).to_host())
            srv.release(t)
        st = srv.stats()
    assert st["completed"] == len(trace) and st["failed"] == 0
    assert st["peak_running"] >= 2, st  # requests overlapped on the device
    identical = 0
    for q, y in zip(trace, got):
        a = _alone(q)
        identical += int(np.array_equal(a, y))
        assert O.rel_l2(y.astype(np.complex128), a.astype(np.complex128)) <= TOL[q.precision], q
    print("bit-identical to alone: %d of %d" % (identical, len(trace)))
    # the oracle on the smallest requests
    for q, y in sorted(zip(trace, got), key=lambda p: p[0].shape[0])[:4]:
        x = O.fill(q.seed, q.shape, np.dtype(q.in_dtype)).astype(
            np.float64 if q.kind == 1 else np.complex128)
        ref = O.rfftn(x) if q.kind == 1 else O.fftn(x)
        e = O.rel_l2(y.astype(np.complex128), ref)
        assert e <= (1e-12 if q.precision == 1 else 1e-5), (q, e)


def test_served_unsnapped_mix_with_bluestein(gpu):
    """The mix with edges rounded to integers (not snapped to 7-smooth
    sizes): requests whose edge has a prime factor above 7 run Bluestein
    passes; served results equal the same transforms alone, and the oracle on
    the smallest."""
    from daggerfft import serve as S
    trace = S.serving_trace(count=24, seed=5, rate=4000.0, lo=16, hi=96, snap=False)
    edges = [q.shape[0] for q in trace]
    print("mix:", [(e, q.kind, q.precision) for e, q in zip(edges, trace)], flush=True)
    assert any(e % p == 0 for e in edges for p in (11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47))
    with S.Server(hbm_budget=4 * S.GiB, max_inflight=8, lanes=4) as srv:
        tickets = [srv.submit(q) for q in trace]
        got = []
        for t, q in zip(tickets, trace):
            assert srv.wait(t, 60.0)
            assert srv.info(t)["status"] == 0
            got.append(srv.result(t, q).to_host())
            srv.release(t)
        st = srv.stats()
    assert st["completed"] == len(trace) and st["failed"] == 0
    for q, y in zip(trace, got):
        a = _alone(q)
        assert O.rel_l2(y.astype(np.complex128), a.astype(np.complex128)) <= TOL[q.precision], q
    for q, y in sorted(zip(trace, got), key=lambda p: p[0].shape[0])[:4]:
        x = O.fill(q.seed, q.shape, np.dtype(q.in_dtype)).astype(
            np.float64 if q.kind == 1 else np.complex128)
        ref = O.rfftn(x) if q.kind == 1 else O.fftn(x)
        assert O.rel_l2(y.astype(np.complex128), ref) <= (1e-12 if q.precision == 1 else 1e-5), q


def test_open_loop_replay_latency_and_results(gpu):
    from daggerfft import serve as S
    trace = _mix(seed=5, count=40, lo=16, hi=128)
    keep = list(range(0, 40, 5))
    with S.Server(hbm_budget=4 * S.GiB, max_inflight=16, lanes=4) as srv:
        res = S.replay(srv, trace, open_loop=True, keep=keep)
        st = srv.stats()
    assert len(res["latencies"]) == 40 and all(x > 0 for x in res["latencies"])
    assert res["latency_p99_ms"] >= res["latency_p50_ms"] > 0
    assert st["plans"] == len(set((q.shape, q.kind, q.precision) for q in trace))
    for r in keep:
        a = _alone(trace[r])
        y = res["results"][r].to_host()
        assert O.rel_l2(y.astype(np.complex128), a.astype(np.complex128)) <= TOL[trace[r].precision], r


def test_budget_admission_queues_and_overtakes(gpu):
    """A budget of a little more than one large request: requests wait for
    memory, small ones overtake a waiting large one, and the server never
    holds more than the budget."""
    from daggerfft import serve as S
    from daggerfft import _lib as L
    big = S.Request((96, 96, 96), L.C2C, L.F64, 77)  # 13.5 MiB: the 14 MiB size class
    small = [S.Request((32, 32, 32), L.C2C, L.F64, 80 + i) for i in range(6)]
    budget = 36 << 20
    with S.Server(hbm_budget=budget, max_inflight=8, lanes=2) as srv:
        order = [big, big, small[0], big, small[1], small[2], big] + small[3:]
        tickets = [srv.submit(q) for q in order]
        outs = []
        for t, q in zip(tickets, order):
            assert srv.wait(t, 60.0)
            assert srv.info(t)["status"] == 0
            outs.append(srv.result(t, q).to_host())
            srv.release(t)
        st = srv.stats()
    assert st["completed"] == len(order)
    assert st["peak_bytes_held"] <= budget, st
    assert st["arena_chunks"] >= 1 and st["buffer_allocs"] >= len(order)
    for q, y in zip(order, outs):
        assert O.rel_l2(y, _alone(q)) <= 1e-12


def test_request_over_budget_is_rejected(gpu):
    from daggerfft import serve as S
    from daggerfft import _lib as L
    with S.Server(hbm_budget=8 << 20, max_inflight=4, lanes=1) as srv:
        with pytest.raises(L.DFFTError) as e:
            srv.submit(S.Request((128, 128, 128), L.C2C, L.F64, 1))
        assert e.value.status == L.DFFT_ERR_BUDGET
        with pytest.raises(L.DFFTError) as e:  # too long for one LDS tile
            srv.submit(S.Request((1, 1, 1 << 15), L.C2C, L.F32, 1))
        assert e.value.status == L.DFFT_ERR_UNSUPPORTED_SIZE


def test_caller_buffers_and_unreleased_results(gpu):
    """Requests that bring their own device input and output; results the
    client has not released keep their memory until it does."""
    import daggerfft as D
    from daggerfft import serve as S
    from daggerfft import _lib as L
    q1 = S.Request((40, 36, 30), L.C2C, L.F64, 11)
    q2 = S.Request((48, 20, 18), L.R2C, L.F32, 12)
    x1 = D.rand(q1.shape, q1.in_dtype, seed=99)
    y1 = D.DeviceArray(q1.out_shape, q1.out_dtype)
    x2 = D.rand(q2.shape, q2.in_dtype, seed=98)
    x1_before = x1.to_host().copy()
    import torch
    torch.cuda.synchronize()  # the server's streams do not wait for the caller's
    with S.Server(hbm_budget=1 * S.GiB, max_inflight=4, lanes=2) as srv:
        t1 = srv.submit(q1, inp=x1, out=y1)
        t2 = srv.submit(q2, inp=x2)
        assert srv.wait(t1, 60.0) and srv.wait(t2, 60.0)
        assert srv.info(t1)["result_ptr"] == y1.ptr
        st = srv.stats()
        assert st["bytes_busy"] > 0  # q2's result is held until released
        r2 = srv.result(t2, q2)
        srv.release(t1)
        srv.release(t2)
        assert srv.stats()["bytes_busy"] == 0
    assert np.array_equal(x1.to_host(), x1_before)  # input left untouched
    assert O.rel_l2(y1.to_host(), D.fft(x1).to_host()) <= 1e-12
    assert O.rel_l2(r2.to_host().astype(np.complex128), D.rfft(x2).to_host().astype(np.complex128)) <= 1e-6


def test_cancel_queued_request(gpu):
    """A result the client holds keeps its memory: with room for one request,
    the next one waits in the queue until it is released or cancelled."""
    from daggerfft import serve as S
    from daggerfft import _lib as L
    q = [S.Request((64, 64, 64), L.C2C, L.F64, 20 + i) for i in range(3)]  # 4 MiB each
    with S.Server(hbm_budget=6 << 20, max_inflight=4, lanes=1) as srv:
        t0 = srv.submit(q[0])
        assert srv.wait(t0, 60.0)
        t1 = srv.submit(q[1])
        t2 = srv.submit(q[2])
        assert not srv.wait(t1, 0.2)  # no memory until t0's result is released
        assert srv.info(t1)["state"] == L.REQ_QUEUED
        srv.release(t1)  # cancel
        with pytest.raises(L.DFFTError):
            srv.info(t1)
        srv.release(t0)
        assert srv.wait(t2, 60.0)
        y2 = srv.result(t2, q[2]).to_host()
        srv.release(t2)
        assert srv.drain(60.0)
        st = srv.stats()
    assert st["completed"] == 2 and st["queued"] == 0
    assert O.rel_l2(y2, _alone(q[2])) <= 1e-12


def test_discarded_results_free_memory_at_completion(gpu):
    """DFFT_REQ_DISCARD: with room for one request's buffer, discarded
    requests run one after another without any release (their buffers return
    to the arena when each completes); their records report no result, a copy
    of it is refused, and a kept request after them is still correct.
    Unknown request flags are rejected."""
    import ctypes

    from daggerfft import serve as S
    from daggerfft import _lib as L
    q = [S.Request((64, 64, 64), L.C2C, L.F64, 30 + i) for i in range(4)]  # 4 MiB each
    with S.Server(hbm_budget=6 << 20, max_inflight=4, lanes=1) as srv:
        ts = [srv.submit(x, discard=True) for x in q[:3]]
        for t in ts:
            assert srv.wait(t, 60.0)
            inf = srv.info(t)
            assert inf["state"] == L.REQ_DONE and inf["result_ptr"] == 0
        assert srv.stats()["bytes_busy"] == 0
        with pytest.raises(L.DFFTError):
            srv.result(ts[0], q[0])
        tk = srv.submit(q[3])
        assert srv.wait(tk, 60.0)
        y = srv.result(tk, q[3]).to_host()
        for t in ts + [tk]:
            srv.release(t)
        r = S._request_struct(q[0], req_flags=0x2)
        t = ctypes.c_int64()
        assert L.lib().dfft_server_submit(srv._h, ctypes.byref(r), ctypes.byref(t)) == L.DFFT_ERR_ARG
    assert O.rel_l2(y, _alone(q[3])) <= 1e-12
