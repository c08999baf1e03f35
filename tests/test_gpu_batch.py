"""Batched solver (many graphs solved as disjoint unions, one per part of
the batch): every graph's kmax, |T| and kmax-truss edge set equal the
single-graph path's; invalid graphs are flagged without disturbing the
others; repeated solves agree; the number of parts changes no answer."""
import os

import numpy as np
import pytest

from conftest import canon_pairs
from graphs import closed_forms

pytestmark = pytest.mark.gpu


def _with_specials(pairs, offs, ns):
    """Append the closed-form graphs (empty, triangle-free, cliques, ...)."""
    ps, os_, nn = [pairs], list(offs), list(ns)
    total = int(offs[-1])
    for name, p, n, _, _ in closed_forms():
        p = canon_pairs(np.asarray(p, dtype=np.uint32).reshape(-1, 2))
        ps.append(p)
        total += len(p)
        os_.append(total)
        nn.append(n)
    return np.concatenate(ps).astype(np.uint32), np.array(os_, dtype=np.uint64), np.array(nn, dtype=np.uint32)


def _solve_batch(ctx, pairs, offs, ns, **kw):
    d = ctx.to_device(np.ascontiguousarray(pairs))
    try:
        return ctx.batch_solve(d, offs, ns, **kw)
    finally:
        d.free()


def _check_against_plain(ctx, pairs, offs, ns, res):
    """Every graph's batched answer equals the plain HIP path's and the
    oracle's (kmax, |T| and, with kept-pair flags, the kmax-truss)."""
    from oracle import oracle as O
    for g in range(len(ns)):
        q = pairs[offs[g]:offs[g + 1]]
        r = ctx.solve_pairs(q, int(ns[g]))
        ref = O.solve_pairs(q, int(ns[g]), want_edges=True)
        assert (r.kmax, r.truss_edges) == (ref.kmax, ref.truss_edges), f"graph {g} n={ns[g]} plain"
        assert (int(res.kmax[g]), int(res.edges[g])) == (ref.kmax, ref.truss_edges), f"graph {g} n={ns[g]}"
        if res.keep is not None:
            kept = q[res.keep[offs[g]:offs[g + 1]] == 1]
            np.testing.assert_array_equal(canon_pairs(kept), ctx.truss_edges(), err_msg=f"graph {g}")
            np.testing.assert_array_equal(canon_pairs(kept), canon_pairs(ref.truss), err_msg=f"graph {g}")


@pytest.mark.parametrize("seed", [11, 12])
def test_batch_matches_plain(ctx, seed):
    from kc_cuda_amd import gen
    pairs, offs, ns = gen.batch(96, 16, 3000, 4.0, 24.0, 3, 20, seed)
    pairs, offs, ns = _with_specials(pairs, offs, ns)
    print(f"[gpu] batch seed={seed} graphs={len(ns)} pairs={len(pairs)}", flush=True)
    res = _solve_batch(ctx, pairs, offs, ns, keep=True)
    assert (res.status == 0).all()
    _check_against_plain(ctx, pairs, offs, ns, res)


def test_batch_large_graphs(ctx):
    """A few graphs at the top of the batch's size range (R-MAT hubs)."""
    from kc_cuda_amd import gen
    pairs, offs, ns = gen.batch(6, 20000, 65536, 16.0, 32.0, 20, 40, 21)
    print(f"[gpu] batch large graphs pairs={len(pairs)} max n={ns.max()}", flush=True)
    res = _solve_batch(ctx, pairs, offs, ns, keep=True)
    assert (res.status == 0).all()
    _check_against_plain(ctx, pairs, offs, ns, res)


def test_batch_flags_invalid_graphs(ctx):
    from kc_cuda_amd import gen
    pairs, offs, ns = gen.batch(8, 16, 400, 6.0, 12.0, 4, 8, 5)
    bad = [np.array([[0, 1], [0, 1]], np.uint32),      # duplicate
           np.array([[2, 1]], np.uint32),               # not canonical
           np.array([[0, 1], [1, 9]], np.uint32),       # id >= n
           np.array([[1, 2], [0, 1]], np.uint32)]       # unsorted
    ps, os_, nn = [pairs], list(offs), list(ns)
    total = int(offs[-1])
    for b in bad:
        ps.append(b)
        total += len(b)
        os_.append(total)
        nn.append(5)
    pairs = np.concatenate(ps).astype(np.uint32)
    offs = np.array(os_, dtype=np.uint64)
    ns = np.array(nn, dtype=np.uint32)
    res = _solve_batch(ctx, pairs, offs, ns, keep=True)
    assert (res.status[:8] == 0).all()
    assert (res.status[8:] == 1).all()
    _check_against_plain(ctx, pairs[:int(offs[8])], offs[:9], ns[:8], res)


def test_batch_repeat_and_autotune(ctx):
    """Repeated solves (with other solves on the context in between) return
    the same answers, non-zero flags are rejected; the autotuner's
    trials agree, its final solve of the whole batch equals the untuned one,
    and so does a solve with the settings it leaves on the context."""
    from kc_cuda_amd import gen
    pairs, offs, ns = gen.batch(64, 32, 5000, 6.0, 30.0, 4, 24, 7)
    d = ctx.to_device(np.ascontiguousarray(pairs))
    try:
        ref = ctx.batch_solve(d, offs, ns, keep=True)
        _check_against_plain(ctx, pairs, offs, ns, ref)
        for _ in range(2):
            r = ctx.batch_solve(d, offs, ns, keep=True)
            np.testing.assert_array_equal(r.kmax, ref.kmax)
            np.testing.assert_array_equal(r.edges, ref.edges)
            np.testing.assert_array_equal(r.keep, ref.keep)
            ctx.solve_pairs(pairs[:int(offs[1])], int(ns[0]))
        from kc_cuda_amd import KtError, _check
        with pytest.raises(KtError, match="sample_every"):
            ctx.batch_autotune(d, offs, ns, sample_every=0)
        run = ctx.batch(d, offs, ns, False)
        try:
            with pytest.raises(KtError, match="reserved"):
                _check(ctx._lib.kt_batch_solve(ctx._h, d.ptr, run.offsets.ctypes.data, run.ns.ctypes.data,
                                               run.count, run.kmax.ptr, run.edges.ptr, run.status.ptr, None, 1))
        finally:
            run.free()
        try:
            for every in (3, 1):
                r, t = ctx.batch_autotune(d, offs, ns, sample_every=every)
                print(f"[gpu] batch autotune every={every}: {t}", flush=True)
                assert 2 + 3 + 3 + 1 + 3 <= t["trials"] <= 2 + 4 + 4 + 1 + 3
                assert t["peel_lanes"] in (0, 8, 16, 32, 64) and t["count_lanes"] in (0, 8, 16, 32, 64)
                assert t["ms_best"] <= t["ms_default"]
                np.testing.assert_array_equal(r.kmax, ref.kmax)
                np.testing.assert_array_equal(r.edges, ref.edges)
                np.testing.assert_array_equal(r.status, ref.status)
                r2 = ctx.batch_solve(d, offs, ns, keep=True)
                np.testing.assert_array_equal(r2.kmax, ref.kmax)
                np.testing.assert_array_equal(r2.edges, ref.edges)
                np.testing.assert_array_equal(r2.keep, ref.keep)
        finally:
            ctx.tune(0, 0, False)
    finally:
        d.free()


def test_batch_empty_and_all_invalid(ctx):
    """A batch of empty graphs and a batch whose graphs are all invalid."""
    d = ctx.to_device(np.zeros((1, 2), np.uint32))
    try:
        r = ctx.batch_solve(d, np.zeros(4, np.uint64), np.array([3, 0, 7], np.uint32), keep=True)
        assert (r.kmax == 0).all() and (r.edges == 0).all() and (r.status == 0).all()
    finally:
        d.free()
    bad = np.array([[2, 1], [0, 0]], np.uint32)
    d = ctx.to_device(bad)
    try:
        r = ctx.batch_solve(d, np.array([0, 1, 2], np.uint64), np.array([3, 3], np.uint32), keep=True)
        assert (r.status == 1).all() and (r.kmax == 0).all() and not r.keep.any()
    finally:
        d.free()


@pytest.mark.parametrize("parts", ["1", "3", "7"])
def test_batch_parts_agree(parts):
    """The batch cut into 1, 3 or 7 concurrent parts (KT_BATCH_PARTS, each
    part's union on its own solver and stream): every graph's kmax, |T|,
    status and kept pairs equal the oracle's, for a mix of R-MAT / ER graphs,
    the closed forms (empty, triangle-free, cliques) and invalid graphs
    spread through the batch, so parts hold invalid and empty graphs."""
    from kc_cuda_amd import Context, gen
    from oracle import oracle as O
    pairs, offs, ns = gen.batch(40, 16, 6000, 6.0, 20.0, 3, 16, 31)
    pairs, offs, ns = _with_specials(pairs, offs, ns)
    # an invalid graph after every 9th valid one
    ps, os_, nn, bad = [], [0], [], []
    total = 0
    for g in range(len(ns)):
        q = pairs[int(offs[g]):int(offs[g + 1])]
        ps.append(q)
        total += len(q)
        os_.append(total)
        nn.append(int(ns[g]))
        if g % 9 == 8:
            ps.append(np.array([[2, 1], [0, 3]], np.uint32))  # not canonical
            total += 2
            os_.append(total)
            nn.append(4)
            bad.append(len(nn) - 1)
    pairs = np.concatenate(ps).astype(np.uint32)
    offs = np.array(os_, dtype=np.uint64)
    ns = np.array(nn, dtype=np.uint32)
    assert len(pairs) >= (1 << 16), "the batch must be large enough to be cut into parts"
    old = os.environ.get("KT_BATCH_PARTS")
    os.environ["KT_BATCH_PARTS"] = parts
    try:
        c = Context(0)  # the batch runner reads the knob when it is made
    finally:
        if old is None:
            os.environ.pop("KT_BATCH_PARTS")
        else:
            os.environ["KT_BATCH_PARTS"] = old
    try:
        print(f"[gpu] batch parts={parts} graphs={len(ns)} pairs={len(pairs)}", flush=True)
        res = _solve_batch(c, pairs, offs, ns, keep=True)
        want_bad = np.zeros(len(ns), dtype=bool)
        want_bad[bad] = True
        np.testing.assert_array_equal(res.status != 0, want_bad)
        for g in range(len(ns)):
            q = pairs[int(offs[g]):int(offs[g + 1])]
            if want_bad[g]:
                assert (int(res.kmax[g]), int(res.edges[g])) == (0, 0)
                assert not res.keep[int(offs[g]):int(offs[g + 1])].any()
                continue
            ref = O.solve_pairs(q, int(ns[g]), want_edges=True)
            assert (int(res.kmax[g]), int(res.edges[g])) == (ref.kmax, ref.truss_edges), f"graph {g}"
            kept = q[res.keep[int(offs[g]):int(offs[g + 1])] == 1]
            np.testing.assert_array_equal(canon_pairs(kept), canon_pairs(ref.truss), err_msg=f"graph {g}")
    finally:
        c.close()
