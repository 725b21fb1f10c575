"""Host logic of the serving loop: arena bookkeeping, request generation, latency percentiles."""
import math

import numpy as np
import pytest

from yeti_amd.serving import ALIGN, ArenaAllocator, ir_session_bytes, latency_summary, request_stream


def test_arena_first_fit_and_coalescing():
    a = ArenaAllocator(10 * ALIGN)
    x = a.alloc(1)               # rounds up to one ALIGN block
    y = a.alloc(3 * ALIGN)
    z = a.alloc(2 * ALIGN + 1)   # three blocks
    assert (x, y, z) == (0, ALIGN, 4 * ALIGN)
    assert a.used == 7 * ALIGN
    assert a.alloc(4 * ALIGN) is None            # only 3 blocks left
    a.release(y, 3 * ALIGN)
    assert a.alloc(4 * ALIGN) is None            # free space is split: 3 + 3
    a.release(x, 1)
    w = a.alloc(4 * ALIGN)                       # [0, 4) coalesced
    assert w == 0
    a.release(w, 4 * ALIGN)
    a.release(z, 2 * ALIGN + 1)
    assert a.free == [(0, 10 * ALIGN)] and a.used == 0
    assert a.peak == 7 * ALIGN


def test_arena_rejects_double_free():
    a = ArenaAllocator(4 * ALIGN)
    x = a.alloc(ALIGN)
    a.release(x, ALIGN)
    with pytest.raises(ValueError):
        a.release(x, ALIGN)
    with pytest.raises(ValueError):
        a.release(3 * ALIGN, 2 * ALIGN)


def test_arena_random_sequence_keeps_invariants():
    rng = np.random.default_rng(0)
    a = ArenaAllocator(1000 * ALIGN)
    live = []
    for _ in range(3000):
        if live and rng.random() < 0.45:
            off, n = live.pop(int(rng.integers(len(live))))
            a.release(off, n)
        else:
            n = int(rng.integers(1, 40 * ALIGN))
            off = a.alloc(n)
            if off is not None:
                live.append((off, n))
        # free list sorted, disjoint, non-adjacent; used + free = size; live ranges disjoint
        fr = a.free
        assert all(fr[i][0] + fr[i][1] < fr[i + 1][0] for i in range(len(fr) - 1))
        assert a.used + sum(l for _, l in fr) == a.size
    spans = sorted((o, o + (n + ALIGN - 1) // ALIGN * ALIGN) for o, n in live)
    assert all(spans[i][1] <= spans[i + 1][0] for i in range(len(spans) - 1))


def test_request_stream_distribution():
    reqs = request_stream(n=4000, rate=100.0, seed=5)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0)
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 0.01) < 0.001          # Poisson, rate 100/s
    nel = np.array([r.nel for r in reqs])
    assert nel.min() >= 4 and nel.max() <= 48
    # log-uniform: the median of elements per direction is near sqrt(4 * 48) ~ 13.9
    assert 12 <= np.median(nel) <= 16
    ps = np.array([r.p for r in reqs])
    assert set(ps.tolist()) == {2, 3, 4}
    assert all(1e3 <= r.E <= 1e5 and 0.2 <= r.nu <= 0.4 and -10 <= r.traction <= -0.1 for r in reqs)
    again = request_stream(n=4000, rate=100.0, seed=5)
    assert [(r.nel, r.p, r.seed) for r in again] == [(r.nel, r.p, r.seed) for r in reqs]


def test_latency_percentiles_nearest_rank():
    s = latency_summary(list(range(1, 101)))
    assert s["p50"] == 50 and s["p99"] == 99 and s["max"] == 100
    assert latency_summary([])["p50"] is None


def test_ir_session_bytes_grows_with_size_and_precision():
    d, d2 = (10, 10, 10), (5, 4, 5)
    assert ir_session_bytes(d, 64) > ir_session_bytes(d, 32) > ir_session_bytes(d2, 32)
    assert ir_session_bytes(d, 32, "fdm") > ir_session_bytes(d, 32, "jacobi")
    assert ir_session_bytes(d, 32) % ALIGN == 0
