"""The allocation-free entry point (kt_solve_pairs_device_noalloc) against
kt_solve_pairs_device on the same device-resident inputs: kmax, |T| and the
kmax-truss edge set must be identical, on two calls in a row (the second
with a smaller graph, which reuses the scratch array the first one left),
then a graph larger than the scratch bound is rejected before any work and
a solve after that still matches."""
import ctypes

import numpy as np
import pytest

from conftest import canon_pairs

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(f"[gpu] {msg}", flush=True)


def _noalloc(ctx, pairs, n, cap):
    from kc_cuda_amd import solve_device_noalloc
    d = ctx.to_device(pairs)
    out = ctx.to_device(np.zeros((max(cap, 1), 2), dtype=np.uint32))
    try:
        r, cnt = solve_device_noalloc(0, d.ptr, len(pairs), n, out.ptr, cap)
        host = np.zeros((max(cap, 1), 2), dtype=np.uint32)
        _check_copy(ctx, host, out)
    finally:
        out.free()
        d.free()
    return r, cnt, canon_pairs(host[:min(cnt, cap)])


def _check_copy(ctx, host, dev):
    from kc_cuda_amd import _check
    _check(ctx._lib.kt_memcpy_d2h(ctx._h, host.ctypes.data, dev.ptr, host.nbytes))


def _plain(ctx, pairs, n):
    d = ctx.to_device(pairs)
    try:
        r = ctx.solve_device(d, len(pairs), n)
    finally:
        d.free()
    return r, ctx.truss_edges()


def test_noalloc_matches_plain(ctx):
    from kc_cuda_amd import gen
    cases = [("rmat15", gen.rmat(15, 16, 0.57, 0.19, 0.19, True, 24, 5)),
             ("er4k_clique12", gen.er_clique(4096, 12.0, 12, 7))]
    for name, (pairs, n) in cases:
        _say(f"noalloc {name}: n={n} m={len(pairs)} plain solve")
        want, want_edges = _plain(ctx, pairs, n)
        _say(f"noalloc {name}: allocation-free solve (kmax={want.kmax} |T|={want.truss_edges})")
        got, cnt, edges = _noalloc(ctx, pairs, n, int(want.truss_edges) + 8)
        assert (got.kmax, got.truss_edges, cnt) == (want.kmax, want.truss_edges, want.truss_edges), name
        assert (got.max_core, got.k_hat, got.m_core) == (want.max_core, want.k_hat, want.m_core), name
        np.testing.assert_array_equal(edges, want_edges, err_msg=name)


def test_noalloc_rejects_oversize_and_recovers(ctx):
    from kc_cuda_amd import KtError, gen, solve_device_noalloc
    pairs, n = gen.er_clique(2048, 10.0, 9, 11)
    d = ctx.to_device(pairs)
    try:
        # 4M edges exceed the 512 MiB scratch bound: rejected on the host
        # before the pointer is read or anything is launched
        with pytest.raises(KtError, match="too large"):
            solve_device_noalloc(0, d.ptr, 4_000_000, 4_000_000)
        _say("noalloc: oversize rejected; small solve after it")
        r, cnt = solve_device_noalloc(0, d.ptr, len(pairs), n)
    finally:
        d.free()
    want, _ = _plain(ctx, pairs, n)
    assert (r.kmax, r.truss_edges, cnt) == (want.kmax, want.truss_edges, want.truss_edges)


def test_noalloc_empty_and_triangle_free(ctx):
    from kc_cuda_amd import solve_device_noalloc
    path = np.array([[0, 1], [1, 2], [2, 3]], dtype=np.uint32)
    d = ctx.to_device(path)
    try:
        r, cnt = solve_device_noalloc(0, d.ptr, len(path), 4)
        assert (r.kmax, r.truss_edges, cnt) == (0, 0, 0)
        r, cnt = solve_device_noalloc(0, d.ptr, 0, 4)
        assert (r.kmax, r.truss_edges, cnt) == (0, 0, 0)
    finally:
        d.free()
