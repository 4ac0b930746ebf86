"""Serving loop: a stream of independent 3D transform requests with plans and
device buffers kept resident (dfft_server_* of include/daggerfft_hip.h).

No reference counterpart as a whole.  It keeps resident what the reference
rebuilds per call: the plan cache (GPU_PLAN_CACHE / get_or_create_gpu_plan,
src/fftgpu.jl:83,107-116, emptied by gpu_cleanup_workspace!,
src/fftgpu.jl:630-634) and the workspace's device staging buffers
(create_gpu_workspace, src/fftgpu.jl:227-249).

Also here: the seeded request trace of the mixed serving workload and the
drivers that replay it open-loop (latency) and saturating (throughput), plus
the two sequential baselines it is compared with.
"""
from __future__ import annotations

import ctypes
import math
import sys
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import DFFTError, check, lib

GiB = 1 << 30


@dataclass(frozen=True)
class Request:
    """One 3D transform request: the input is the uniform[0,1) generator
    stream of `seed` (real for R2C, complex for C2C); `arrival` is seconds
    after the start of the trace."""
    shape: tuple
    kind: int          # L.C2C or L.R2C
    precision: int     # L.F32 or L.F64
    seed: int
    arrival: float = 0.0
    sign: int = -1

    @property
    def in_dtype(self):
        if self.kind == L.R2C:
            return np.float64 if self.precision == L.F64 else np.float32
        return np.complex128 if self.precision == L.F64 else np.complex64

    @property
    def out_shape(self):
        if self.kind == L.R2C:
            return (self.shape[0] // 2 + 1,) + tuple(self.shape[1:])
        return tuple(self.shape)

    @property
    def out_dtype(self):
        return np.complex128 if self.precision == L.F64 else np.complex64

    def algorithmic_bytes(self):
        """Bytes read + written over the three axis passes."""
        n = int(np.prod(self.shape))
        es = 8 if self.precision == L.F64 else 4
        if self.kind == L.R2C:
            nc = int(np.prod(self.out_shape))
            return n * es + nc * 2 * es + 2 * (2 * nc * 2 * es)
        return 3 * 2 * n * 2 * es


def smooth_even_sizes(lo=64, hi=512):
    """Even integers in [lo, hi] with no prime factor above 7."""
    out = []
    for m in range(lo + (lo & 1), hi + 1, 2):
        r = m
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        if r == 1:
            out.append(m)
    return out


def snap_size(x, sizes):
    """The size in `sizes` nearest to x (the smaller one on a tie)."""
    i = int(np.searchsorted(sizes, x))
    if i <= 0:
        return sizes[0]
    if i >= len(sizes):
        return sizes[-1]
    a, b = sizes[i - 1], sizes[i]
    return a if x - a <= b - x else b


def serving_trace(count=2000, seed=8, rate=200.0, lo=64, hi=512, c2c_frac=0.6, f32_frac=0.5, data_seed0=10000,
                  snap=True):
    """The mixed serving workload.  Request r draws u_a = u(seed, 4r + a) from
    the SplitMix64 counter generator:
      inter-arrival  -ln(1 - u_0) / rate seconds (Poisson arrivals),
      kind           C2C if u_1 < c2c_frac else R2C,
      edge           lo * (hi/lo)^u_2 (log-uniform), snapped to the nearest
                     even {2,3,5,7}-smooth size (snap=False: rounded to the
                     nearest integer, so edges with a prime factor above 7 run
                     Bluestein passes); the request is edge^3,
      precision      F32 if u_3 < f32_frac else F64,
    and its input is the generator stream of seed data_seed0 + r."""
    from . import _splitmix_u
    sizes = smooth_even_sizes(lo, hi)
    t = 0.0
    out = []
    for r in range(count):
        u = [_splitmix_u(seed, 4 * r + a) for a in range(4)]
        t += -math.log1p(-u[0]) / rate
        kind = L.C2C if u[1] < c2c_frac else L.R2C
        x = lo * (hi / lo) ** u[2]
        edge = snap_size(x, sizes) if snap else int(round(x))
        prec = L.F32 if u[3] < f32_frac else L.F64
        out.append(Request((edge, edge, edge), kind, prec, data_seed0 + r, t))
    return out


def _request_struct(req, t_arrival=0.0, d_in=None, d_out=None, decomp=L.PENCIL, flags=0, req_flags=0):
    r = L.dfft_request()
    r.n[0], r.n[1], r.n[2] = [int(s) for s in req.shape]
    r.kind = req.kind
    r.sign = req.sign
    r.precision = req.precision
    r.decomp = decomp
    r.flags = flags
    r.req_flags = req_flags
    r.seed = req.seed
    r.d_in = d_in
    r.d_out = d_out
    r.t_arrival = t_arrival
    return r


class Server:
    """A serving loop on the current device.

    hbm_budget: device bytes the server may hold for request buffers and plans;
    max_inflight: requests admitted and not yet retired; lanes: HIP streams."""

    def __init__(self, hbm_budget=200 * GiB, max_inflight=32, lanes=4):
        from . import _torch
        _torch().cuda.init()
        cfg = L.dfft_server_config(int(hbm_budget), int(max_inflight), int(lanes), 0, 0)
        h = ctypes.c_void_p()
        check(lib().dfft_server_create(ctypes.byref(h), ctypes.byref(cfg)), "dfft_server_create")
        self._h = h

    def clock(self):
        t = ctypes.c_double()
        check(lib().dfft_server_clock(self._h, ctypes.byref(t)))
        return t.value

    def prepare(self, reqs):
        """Build and keep the plans of these request shapes."""
        reqs = list(reqs)
        arr = (L.dfft_request * max(len(reqs), 1))(*[_request_struct(q) for q in reqs])
        check(lib().dfft_server_prepare(self._h, arr, len(reqs)), "dfft_server_prepare")

    def submit(self, req, t_arrival=0.0, inp=None, out=None, discard=False):
        """Queue `req`; returns its ticket.  inp/out: optional DeviceArrays the
        request reads / writes instead of server buffers (kept alive by the
        caller).  discard: the result is not read (DFFT_REQ_DISCARD): its
        server buffer returns to the arena when the request completes."""
        r = _request_struct(req, t_arrival, inp.ptr if inp is not None else None,
                            out.ptr if out is not None else None, req_flags=L.REQ_DISCARD if discard else 0)
        t = ctypes.c_int64()
        check(lib().dfft_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)), "dfft_server_submit")
        return t.value

    def wait(self, ticket, timeout=-1.0):
        st = lib().dfft_server_wait(self._h, int(ticket), float(timeout))
        if st == L.DFFT_ERR_NOT_READY:
            return False
        check(st, "dfft_server_wait")
        return True

    def wait_any(self, timeout=-1.0):
        """The next finished ticket (completion order), or None on time-out."""
        t = ctypes.c_int64()
        st = lib().dfft_server_wait_any(self._h, float(timeout), ctypes.byref(t))
        if st == L.DFFT_ERR_NOT_READY:
            return None
        check(st, "dfft_server_wait_any")
        return t.value

    def info(self, ticket):
        i = L.dfft_request_info()
        check(lib().dfft_server_info(self._h, int(ticket), ctypes.byref(i)), "dfft_server_info")
        return {"state": i.state, "status": i.status, "lane": i.lane, "t_arrival": i.t_arrival,
                "t_admit": i.t_admit, "t_done": i.t_done, "result_ptr": i.d_result or 0,
                "result_bytes": i.result_bytes, "algorithmic_bytes": i.algorithmic_bytes}

    def result(self, ticket, req, stream=None):
        """A copy of a finished request's result as a DeviceArray."""
        from . import DeviceArray, _stream_handle
        out = DeviceArray(req.out_shape, req.out_dtype)
        st = lib().dfft_server_copy_result(self._h, int(ticket), ctypes.c_void_p(out.ptr), out.nbytes,
                                           _stream_handle(stream))
        check(st, "dfft_server_copy_result")
        return out

    def release(self, ticket):
        check(lib().dfft_server_release(self._h, int(ticket)), "dfft_server_release")

    def drain(self, timeout=-1.0):
        st = lib().dfft_server_drain(self._h, float(timeout))
        if st == L.DFFT_ERR_NOT_READY:
            return False
        check(st, "dfft_server_drain")
        return True

    def stats(self):
        s = L.dfft_server_stats()
        check(lib().dfft_server_stats_get(self._h, ctypes.byref(s)), "dfft_server_stats_get")
        return {k: getattr(s, k) for k, _ in L.dfft_server_stats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().dfft_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _percentile(xs, q):
    if not xs:
        return float("nan")
    return float(np.percentile(np.asarray(xs), q))


def replay(server, trace, open_loop=True, keep=None, time_scale=1.0):
    """Replays `trace` through `server`.

    open_loop: request r is submitted at its arrival time (trace[r].arrival *
    time_scale after the start) whatever the state of earlier ones, and its
    latency is t_done - scheduled arrival.  Otherwise every request is
    submitted at once (saturating) and the run measures throughput.
    keep: indices of requests whose results are copied out before release.
    Returns a dict with per-request latencies, throughput and kept results."""
    keep = set(keep or ())
    n = len(trace)
    index = {}
    lat = [None] * n
    kept = {}
    errors = []
    lock = threading.Lock()
    done_count = [0]
    expected = [n]  # lowered if submission stops early

    def reaper():
        while True:
            with lock:
                if done_count[0] >= expected[0]:
                    return
            t = server.wait_any(0.5)
            if t is None:
                continue
            while True:  # the ticket may finish before submit() has returned it
                with lock:
                    r = index.get(t)
                if r is not None:
                    break
                time.sleep(0)
            inf = server.info(t)
            if inf["state"] != L.REQ_DONE:
                errors.append((r, inf["status"]))
            else:
                lat[r] = inf["t_done"] - inf["t_arrival"]
                if r in keep:
                    kept[r] = server.result(t, trace[r])
            server.release(t)
            with lock:
                done_count[0] += 1

    # The submitter and the reaper share the interpreter lock.  With the
    # default 5 ms switch interval a saturating submit loop keeps the reaper
    # from releasing finished results for milliseconds at a time, so they pile
    # up in the server's memory; a 0.1 ms interval interleaves the two.
    switch = sys.getswitchinterval()
    sys.setswitchinterval(1e-4)
    th = threading.Thread(target=reaper, daemon=True)
    th.start()
    # open loop: the first arrival is 50 ms ahead, so the submitter is ready for it
    t0 = server.clock() + (0.05 if open_loop else 0.0)
    submitted = 0
    try:
        for r, q in enumerate(trace):
            due = t0 + (q.arrival * time_scale if open_loop else 0.0)
            if open_loop:
                while True:
                    now = server.clock()
                    if now >= due:
                        break
                    if due - now > 0.002:
                        time.sleep(due - now - 0.001)
            # results outside `keep` are never read: the server frees them at
            # completion instead of waiting for the reaper's release
            t = server.submit(q, t_arrival=due if open_loop else 0.0, discard=r not in keep)
            with lock:
                index[t] = r
            submitted += 1
    finally:
        with lock:
            expected[0] = submitted
        th.join()
        sys.setswitchinterval(switch)
    t_end = server.clock()
    if errors:
        raise DFFTError(errors[0][1], "serving request %d" % errors[0][0])
    from . import _torch
    _torch().cuda.synchronize()
    span = t_end - t0
    total_bytes = sum(q.algorithmic_bytes() for q in trace)
    lats = [x for x in lat if x is not None]
    return {"requests": n, "seconds": span, "req_per_s": n / span, "GBps": total_bytes / span / 1e9,
            "latency_p50_ms": 1e3 * _percentile(lats, 50), "latency_p99_ms": 1e3 * _percentile(lats, 99),
            "latency_mean_ms": 1e3 * float(np.mean(lats)), "latencies": lats, "results": kept}


def sequential_plain(trace, keep=None, cold=False):
    """The same requests as plain one-at-a-time library calls, each waited for
    before the next: x = rand(shape, seed); y = fft(x) (or rfft); synchronize.

    warm (cold=False): the plain API's plan cache and torch's caching allocator
    keep plans and buffers between calls, as GPU_PLAN_CACHE and CUDA.jl's pool
    do for the reference.  cold=True: every call builds its plan and allocates
    its arrays, and both are freed after it (clear_plan_cache, which is
    gpu_cleanup_workspace!'s emptying of GPU_PLAN_CACHE, src/fftgpu.jl:630-634,
    and torch.cuda.empty_cache)."""
    from . import _torch, clear_plan_cache, fft, rand, rfft
    torch = _torch()
    keep = set(keep or ())
    kept = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r, q in enumerate(trace):
        x = rand(q.shape, q.in_dtype, seed=q.seed)
        y = rfft(x) if q.kind == L.R2C else fft(x)
        torch.cuda.synchronize()
        if r in keep:
            kept[r] = y
        del x, y
        if cold:
            clear_plan_cache()
            torch.cuda.empty_cache()
    span = time.perf_counter() - t0
    total_bytes = sum(q.algorithmic_bytes() for q in trace)
    return {"requests": len(trace), "seconds": span, "req_per_s": len(trace) / span,
            "GBps": total_bytes / span / 1e9, "results": kept}
