"""GPU parity of the reference-compatible stage functions (kc_gpu.h / C-ABI
twins in kc_gpu_c.h) against the oracle's restatement of each stage, and of
the reference's own host driver (src/kc_gpu.cpp compiled from its sources)
linked against libkc_amd.so."""
import json

import numpy as np
import pytest

from conftest import GOLDEN
from graphs import random_graph

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


def _graph(kind):
    from kc_cuda_amd import gen
    if kind == "random":
        return random_graph(400, 0.05, 17, cliques=(11, 8)), 400
    if kind == "rmat12":
        return gen.rmat(12, 16, 0.57, 0.19, 0.19, True, 0, 1)
    return gen.er_clique(3000, 10.0, 17, 11)


def _csr32(pairs, n):
    from oracle import oracle as O
    off, adj, n = O.csr_from_pairs(pairs, n)
    return off.astype(np.uint32), adj, n, np.diff(off).astype(np.uint32)


@pytest.mark.parametrize("kind", ["random", "rmat12", "er3k"])
def test_stage_functions_match_oracle(kind):
    from kc_cuda_amd import compat as C
    from oracle import oracle as O
    pairs, n = _graph(kind)
    print(f"[gpu] compat stages {kind} m={len(pairs)}", flush=True)
    nbr_u, nbr, n, deg = _csr32(pairs, n)
    off64 = nbr_u.astype(np.uint64)
    # (b1) max-core and a fixed-k core
    mask = np.zeros(n, dtype=np.uint32)
    r = C.kmax_core_gpu(n, nbr_u, nbr, deg, mask, NONE)
    cmax, removed = O.max_core(off64, nbr, n)
    assert r == cmax - 1
    np.testing.assert_array_equal(mask, removed.astype(np.uint32))
    L = max(3, cmax // 2)
    mask2 = np.zeros(n, dtype=np.uint32)
    assert C.kmax_core_gpu(n, nbr_u, nbr, deg, mask2, L) == L - 1
    np.testing.assert_array_equal(mask2, O.kcore_removed(off64, nbr, n, L).astype(np.uint32))
    # (a1) ranks and oriented degrees on the max-core
    dwm = np.zeros(n, dtype=np.uint32)
    to = np.zeros(n, dtype=np.uint32)
    li, new_n, cdeg, ddir = C.count_deg_with_mask(nbr, nbr_u, n, mask, dwm, to)
    nc, oto, inv, degc = O.rank(off64, nbr, n, removed)
    assert new_n == nc
    np.testing.assert_array_equal(li[:, 0], inv)
    np.testing.assert_array_equal(li[:, 2], degc[inv])
    np.testing.assert_array_equal(cdeg, degc[inv])
    off_d, adj_d = O.oriented(off64, nbr, n, oto, nc)
    np.testing.assert_array_equal(ddir, np.diff(off_d).astype(np.uint32))
    np.testing.assert_array_equal(to[inv], np.arange(nc, dtype=np.uint32))
    # (a2) oriented CSR
    nud = off_d.astype(np.uint32)
    nd = C.build_nbr_direct(nbr, nbr_u, n, mask, to, nud, new_n)
    np.testing.assert_array_equal(nd, adj_d)
    # (a3) undirected CSR with edge ids and sons
    unbr, fa, ls, rs = C.build_graph_from_direct(nd, nud, new_n)
    off_u, adj_u, ofa, lson, rson = O.undirected(off_d, adj_d, nc)
    np.testing.assert_array_equal(unbr, adj_u)
    np.testing.assert_array_equal(fa, ofa.astype(np.uint32))
    np.testing.assert_array_equal(ls, lson.astype(np.uint32))
    np.testing.assert_array_equal(rs, rson.astype(np.uint32))
    # (c1) doubled supports and their per-slot copies
    tc, tcd = C.tricount_direct(nud, nd, new_n, ls, rs)
    sup2 = O.support2(off_d, adj_d, nc)
    np.testing.assert_array_equal(tc, sup2)
    np.testing.assert_array_equal(tcd[ls], sup2)
    np.testing.assert_array_equal(tcd[rs], sup2)
    tc_lim, _ = C.tricount_direct(nud, nd, new_n, tc_limit=5)
    np.testing.assert_array_equal(tc_lim, sup2)


@pytest.mark.parametrize("kind", ["random", "rmat12", "er3k"])
def test_two_pass_driver_matches_oracle(kind):
    from kc_cuda_amd import compat as C
    from oracle import oracle as O
    pairs, n = _graph(kind)
    nbr_u, nbr, n, deg = _csr32(pairs, n)
    print(f"[gpu] compat two-pass {kind}", flush=True)
    ref = O.solve_pairs(pairs, n)
    assert C.solve(n, nbr_u, nbr, deg) == (ref.kmax, ref.truss_edges)


def test_kmax_truss_gpu_levels():
    """kmax_truss_GPU searches the doubled thresholds above k (first level
    k + 2 k_step, early exit when it keeps `ans` edges, back-off below)."""
    from kc_cuda_amd import compat as C
    from oracle import oracle as O
    pairs, n = _graph("random")
    off, adj, n = O.csr_from_pairs(pairs, n)
    nc, to, inv, _ = O.rank(off, adj, n, np.zeros(n, dtype=np.uint8))
    off_d, adj_d = O.oriented(off, adj, n, to, nc)
    ref = O.solve_pairs(pairs, n)
    nud = off_d.astype(np.uint32)
    # plain search from the bottom
    assert C.kmax_truss_GPU(0, nc, 0, 19, nud, adj_d) == (ref.kmax, ref.truss_edges)
    # first level above kmax: backs off to the answer
    assert C.kmax_truss_GPU(0, nc, 0, ref.kmax + 3, nud, adj_d) == (ref.kmax, ref.truss_edges)
    # nothing reachable above the answer: outputs unchanged
    k = 2 * (ref.kmax - 2)
    assert C.kmax_truss_GPU(0, nc, k, 1, nud, adj_d, 7, 9) == (7, 9)
    # early exit at the first level when its size equals `ans`
    lvl3 = C.kmax_truss_GPU(0, nc, 0, 1, nud, adj_d)
    k3 = O.solve_pairs(pairs, n)  # noqa: F841
    first = C.kmax_truss_GPU(0, nc, 2 * (3 - 2) - 2, 1, nud, adj_d)
    assert first == lvl3


def test_reference_driver_linked_against_library(tmp_path):
    """src/kc_gpu.cpp, built from the reference's sources into oracle/_ref
    and linked against libkc_amd.so, gives the reference's answers."""
    from kc_cuda_amd import gen
    from oracle import oracle as O
    if not O.KC_GPU_LINKED.exists():
        pytest.skip("oracle/_ref/kc_gpu_linked not built (needs the reference sources)")
    gold = json.loads((GOLDEN / "kc_cpu_results.json").read_text())
    for name in ("er3k_clique17", "rmat12", "rmat14_clique12_s7"):
        from graphs import make_workload
        pairs, n = make_workload(gold[name]["params"])
        path = tmp_path / f"{name}.tsv"
        gen.write_tsv(path, pairs)
        print(f"[gpu] kc_gpu_linked {name}", flush=True)
        r = O.run_kc_cpu(path, threads=4, timeout=120, binary=O.KC_GPU_LINKED)
        assert (r.kmax, r.truss_edges) == (gold[name]["kmax"], gold[name]["truss_edges"])
    exp = json.loads((GOLDEN / "tiny" / "expected.json").read_text())
    for fname, line in list(exp.items())[:6]:
        r = O.run_kc_cpu(GOLDEN / "tiny" / fname, threads=2, timeout=60, binary=O.KC_GPU_LINKED)
        assert f"kmax = {r.kmax}, Edges in kmax-truss = {r.truss_edges}." == line


def test_long_rows_clique_k8250(ctx):
    """K_8250: every oriented row but the last is longer than 8192 entries
    before it shrinks, so build_nbr_direct takes the segmented-sort branch;
    closed forms: rank r = 8249 - id (equal degrees, id descending), row r
    holds ranks r+1..8249, every doubled support is 2 * 8248, and the whole
    clique is its kmax-truss: (8250, 34,027,125). The native solve sees the
    same graph as a dense core of 8250 vertices."""
    from kc_cuda_amd import compat as C
    n = 8250
    m = n * (n - 1) // 2
    u, v = np.triu_indices(n, 1)
    pairs = np.stack([u, v], axis=1).astype(np.uint32)
    del u, v
    print(f"[gpu] K{n}: native solve, m={m}", flush=True)
    r = ctx.solve_pairs(pairs, n)
    assert (r.kmax, r.truss_edges) == (n, m)
    assert r.max_core == n - 1 and r.n_core == n and r.m_core == m
    # the reference's CSR: row y holds every other vertex, ascending
    nbr_u = (np.arange(n + 1, dtype=np.uint64) * (n - 1)).astype(np.uint32)
    full = np.tile(np.arange(n, dtype=np.uint32), n).reshape(n, n)
    nbr = full[~np.eye(n, dtype=bool)].astype(np.uint32)
    del full
    deg = np.full(n, n - 1, dtype=np.uint32)
    print(f"[gpu] K{n}: compat stages", flush=True)
    mask = np.zeros(n, dtype=np.uint32)
    assert C.kmax_core_gpu(n, nbr_u, nbr, deg, mask, NONE) == n - 2
    assert not mask.any()
    dwm = np.zeros(n, dtype=np.uint32)
    to = np.zeros(n, dtype=np.uint32)
    li, new_n, cdeg, ddir = C.count_deg_with_mask(nbr, nbr_u, n, mask, dwm, to)
    assert new_n == n
    np.testing.assert_array_equal(to, (n - 1 - np.arange(n)).astype(np.uint32))
    np.testing.assert_array_equal(ddir, (n - 1 - np.arange(n)).astype(np.uint32))
    nud = np.concatenate([[0], np.cumsum(ddir, dtype=np.uint64)]).astype(np.uint32)
    nd = C.build_nbr_direct(nbr, nbr_u, n, mask, to, nud, new_n)
    want = np.concatenate([np.arange(rk + 1, n, dtype=np.uint32) for rk in range(n)])
    np.testing.assert_array_equal(nd, want)
    del want
    unbr, fa, ls, rs = C.build_graph_from_direct(nd, nud, new_n)
    tc, tcd = C.tricount_direct(nud, nd, new_n, ls, rs)
    assert tc.size == m and (tc == 2 * (n - 2)).all()
    assert (tcd == 2 * (n - 2)).all()
    print(f"[gpu] K{n}: compat two-pass driver", flush=True)
    assert C.solve(n, nbr_u, nbr, deg) == (n, m)
