"""Serving loop: every query's kmax and |T| equal its stand-alone solve when it
runs in a mix of concurrent queries; admission keeps the bytes charged to the
slots within the HBM budget, rejects what can never fit, and no slot grows
past what its query was charged (the footprint model is an upper bound)."""
import numpy as np
import pytest

from conftest import canon_pairs
from graphs import closed_forms

pytestmark = pytest.mark.gpu


def _mix(seed: int, count: int, median: int, sigma: float, cap: int):
    from kc_cuda_amd import gen
    qs = gen.serve_mix(count, median, sigma, cap, seed)
    for _, p, n, _, _ in closed_forms():
        qs.append((canon_pairs(np.asarray(p, dtype=np.uint32).reshape(-1, 2)), n))
    return qs


def _alone(ctx, qs):
    out = []
    for p, n in qs:
        r = ctx.solve_pairs(p, n)
        out.append((r.kmax, r.truss_edges))
    return out


@pytest.mark.parametrize("slots", [1, 4, 8])
def test_serve_results_equal_alone(ctx, slots):
    from kc_cuda_amd import Server
    qs = _mix(41, 24, 60_000, 1.2, 2_000_000)
    print(f"[gpu] serve mix slots={slots} queries={len(qs)} edges={sum(len(p) for p, _ in qs)}", flush=True)
    ref = _alone(ctx, qs)
    with Server(0, slots, 32 << 30) as srv:
        tickets = [srv.submit(p, n) for p, n in qs]
        res = [srv.wait(t) for t in tickets]
        c = srv.counters()
    for i, (r, (k, e)) in enumerate(zip(res, ref)):
        assert r.status == 0, r.error
        assert (r.kmax, r.truss_edges) == (k, e), f"query {i}"
        assert r.slot_bytes <= r.charged, f"query {i}: slot holds {r.slot_bytes} > charged {r.charged}"
    assert c["completed"] == len(qs) and c["failed"] == 0 and c["over_estimate"] == 0
    assert c["peak_in_flight"] <= slots
    assert c["in_flight"] == 0 and c["bytes_in_use"] <= c["peak_bytes_in_use"]


def test_serve_admission_by_footprint(ctx):
    """A budget that fits about two of the largest queries: the charged bytes
    never exceed it, and a query bigger than the whole budget is rejected
    without disturbing the others."""
    from kc_cuda_amd import Server, gen
    qs = _mix(42, 16, 100_000, 1.0, 1_000_000)
    big = max(Server.footprint(len(p), n) for p, n in qs)
    budget = 2 * big + (1 << 20)
    huge, hn = gen.serve_query(6_000_000, "chung_lu", 20, 9)
    assert Server.footprint(len(huge), hn) > budget
    ref = _alone(ctx, qs)
    with Server(0, 8, budget) as srv:
        t_huge = srv.submit(huge, hn)
        tickets = [srv.submit(p, n) for p, n in qs]
        res = [srv.wait(t) for t in tickets]
        r_huge = srv.wait(t_huge)
        c = srv.counters()
    assert r_huge.status == 1 and "budget" in r_huge.error
    for i, (r, (k, e)) in enumerate(zip(res, ref)):
        assert r.status == 0, r.error
        assert (r.kmax, r.truss_edges) == (k, e), f"query {i}"
    assert c["rejected"] == 1 and c["completed"] == len(qs)
    assert c["peak_bytes_in_use"] <= budget
    assert c["over_estimate"] == 0


def test_serve_device_resident_inputs(ctx):
    from kc_cuda_amd import Server
    qs = _mix(43, 8, 50_000, 1.0, 500_000)
    ref = _alone(ctx, qs)
    bufs = [ctx.to_device(np.ascontiguousarray(p)) for p, _ in qs]
    try:
        with Server(0, 4, 16 << 30) as srv:
            tickets = [srv.submit(b, n) for b, (_, n) in zip(bufs, qs)]
            res = [srv.wait(t) for t in tickets]
    finally:
        for b in bufs:
            b.free()
    for i, (r, (k, e)) in enumerate(zip(res, ref)):
        assert r.status == 0, r.error
        assert (r.kmax, r.truss_edges) == (k, e), f"query {i}"
