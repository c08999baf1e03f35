"""Serving loop: every query's kmax and |T| equal its stand-alone solve when it
runs in a mix of concurrent queries, grouped into a union with other queries
or not; admission keeps the bytes charged to the slots within the HBM budget,
rejects what can never fit, and no slot grows past what its queries were
charged (the footprint model is an upper bound)."""
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
    """Stand-alone plain solves, each checked against the oracle."""
    from oracle import oracle as O
    out = []
    for i, (p, n) in enumerate(qs):
        r = ctx.solve_pairs(p, n)
        ref = O.solve_pairs(p, n)
        assert (r.kmax, r.truss_edges) == (ref.kmax, ref.truss_edges), f"query {i} alone vs oracle"
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
    assert c["in_flight"] == 0 and c["bytes_in_use"] <= c["peak_bytes_in_use"]
    assert c["grouped"] == sum(r.group > 1 for r in res)
    if slots == 1:  # everything behind the first query waits: small ones are grouped
        assert c["grouped"] > 0
        assert c["peak_in_flight"] > 1


def test_serve_group_falls_back_per_query(ctx):
    """A group that holds inputs the batched solve does not take (pairs out of
    order, pairs with u > v, an id >= n) solves those queries alone: they get
    their stand-alone answers or error, the rest of the group its union
    answers."""
    from kc_cuda_amd import Server, gen
    rng = np.random.default_rng(5)
    good = [gen.serve_query(t, "chung_lu", 12, s) for t, s in ((40_000, 1), (60_000, 2), (20_000, 3))]
    shuffled, sn = gen.serve_query(50_000, "chung_lu", 14, 4)
    shuffled = np.ascontiguousarray(shuffled[rng.permutation(len(shuffled))])
    flipped, fn = gen.serve_query(30_000, "rmat", 10, 5)
    flipped = np.ascontiguousarray(flipped[:, ::-1])
    bad, _ = gen.serve_query(30_000, "chung_lu", 10, 6)
    empty = (np.zeros((0, 2), np.uint32), 10)
    qs = [good[0], (shuffled, sn), good[1], (flipped, fn), (bad, int(bad.max())), empty, good[2]]
    bad_i = 4
    assert all(len(p) <= 1 << 16 for p, _ in qs)  # each small enough to join a group
    ref = {i: (r.kmax, r.truss_edges) for i, r in
           ((i, ctx.solve_pairs(p, n)) for i, (p, n) in enumerate(qs) if i != bad_i)}
    with Server(0, 1, 16 << 30) as srv:
        blocker, blk_n = gen.serve_query(3_000_000, "chung_lu", 20, 7)
        tb = srv.submit(blocker, blk_n)  # too big to group; holds the slot while the rest queue up
        tickets = [srv.submit(p, n) for p, n in qs]
        res = [srv.wait(t) for t in tickets]
        rb = srv.wait(tb)
        c = srv.counters()
    assert rb.status == 0 and rb.group == 1
    for i, r in enumerate(res):
        if i == bad_i:
            assert r.status == 2 and "vertex id" in r.error
            continue
        assert r.status == 0, (i, r.error)
        assert (r.kmax, r.truss_edges) == ref[i], f"query {i}"
    assert {i for i, r in enumerate(res) if r.group > 1} == {0, 2, 5, 6}
    assert c["grouped"] == 4 and c["failed"] == 1


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
