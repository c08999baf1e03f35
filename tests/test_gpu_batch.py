"""Batched solver (many graphs solved as one disjoint union): every graph's
kmax, |T| and kmax-truss edge set equal the single-graph path's; invalid
graphs are flagged without disturbing the others; repeated solves agree."""
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
