"""Host-side logic of the serving workload: the seeded request trace, the
size snapping and the per-request byte counts (no GPU)."""
import math

import numpy as np

from oracle import oracle as O


def _smooth(m):
    for p in (2, 3, 5, 7):
        while m % p == 0:
            m //= p
    return m == 1


def test_smooth_even_sizes_brute_force():
    from daggerfft import serve as S
    sizes = S.smooth_even_sizes(64, 512)
    assert sizes == [m for m in range(64, 513) if m % 2 == 0 and _smooth(m)]
    assert sizes[0] == 64 and sizes[-1] == 512 and len(sizes) == 55


def test_snap_size_nearest_and_ties():
    from daggerfft import serve as S
    sizes = [64, 70, 72, 80]
    assert S.snap_size(63.0, sizes) == 64
    assert S.snap_size(66.9, sizes) == 64
    assert S.snap_size(67.0, sizes) == 64  # tie -> smaller
    assert S.snap_size(67.1, sizes) == 70
    assert S.snap_size(76.0, sizes) == 72  # tie -> smaller
    assert S.snap_size(1e9, sizes) == 80


def test_trace_draws_from_the_counter_generator():
    """Request r's four draws are u(seed, 4r + a), the same stream the C
    oracle's generator produces."""
    from daggerfft import serve as S
    from daggerfft import _lib as L
    tr = S.serving_trace(count=6, seed=8)
    t = 0.0
    sizes = S.smooth_even_sizes(64, 512)
    for r, q in enumerate(tr):
        u = [O.uniform(8, 4 * r + a) for a in range(4)]
        t += -math.log(1.0 - u[0]) / 200.0
        assert abs(q.arrival - t) < 1e-12
        assert q.kind == (L.C2C if u[1] < 0.6 else L.R2C)
        e = 64 * 8.0 ** u[2]
        assert q.shape == (S.snap_size(e, sizes),) * 3
        assert abs(q.shape[0] - e) <= max(b - a for a, b in zip(sizes, sizes[1:])) / 2 + 1e-9
        assert q.precision == (L.F32 if u[3] < 0.5 else L.F64)
        assert q.seed == 10000 + r


def test_trace_distribution():
    from daggerfft import serve as S
    from daggerfft import _lib as L
    tr = S.serving_trace()
    assert len(tr) == 2000
    arr = np.array([q.arrival for q in tr])
    assert np.all(np.diff(arr) > 0)
    assert abs(arr[-1] / 2000 - 1 / 200) < 0.1 / 200  # mean inter-arrival 5 ms
    assert abs(np.mean([q.kind == L.C2C for q in tr]) - 0.6) < 0.035
    assert abs(np.mean([q.precision == L.F32 for q in tr]) - 0.5) < 0.035
    e = np.array([q.shape[0] for q in tr])
    assert e.min() >= 64 and e.max() <= 512
    assert abs(np.median(e) - 64 * 8 ** 0.5) < 15  # log-uniform: median = geometric mean
    assert abs(np.mean(np.log(e / 64) / np.log(8)) - 0.5) < 0.03
    assert S.serving_trace(count=50) == tr[:50]  # deterministic


def test_request_algorithmic_bytes():
    from daggerfft import serve as S
    from daggerfft import _lib as L
    n = 64 ** 3
    assert S.Request((64, 64, 64), L.C2C, L.F64, 1).algorithmic_bytes() == 3 * 2 * n * 16
    assert S.Request((64, 64, 64), L.C2C, L.F32, 1).algorithmic_bytes() == 3 * 2 * n * 8
    nc = 33 * 64 * 64
    # R2C: real in + half-spectrum out, then two complex passes over the half spectrum
    assert S.Request((64, 64, 64), L.R2C, L.F64, 1).algorithmic_bytes() == n * 8 + nc * 16 + 4 * nc * 16
    assert S.Request((64, 64, 64), L.R2C, L.F32, 1).out_shape == (33, 64, 64)


def test_unsnapped_trace_keeps_integer_edges():
    """snap=False rounds the same log-uniform draws to integers: the arrivals,
    kinds and precisions are those of the snapped trace, and edges with prime
    factors above 7 appear (they run Bluestein passes)."""
    from daggerfft import serve as S
    a = S.serving_trace(count=300, seed=8)
    b = S.serving_trace(count=300, seed=8, snap=False)
    assert [(q.kind, q.precision, q.arrival) for q in a] == [(q.kind, q.precision, q.arrival) for q in b]
    edges = [q.shape[0] for q in b]
    assert all(64 <= e <= 512 for e in edges)
    assert sum(not _smooth(e) for e in edges) > 100
    for qa, qb in zip(a, b):
        assert abs(qa.shape[0] - qb.shape[0]) <= max(8, 0.06 * qb.shape[0])
