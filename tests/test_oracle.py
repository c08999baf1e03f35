"""Pins the oracle (oracle/kt_oracle.c) before anything is checked against it.

Each function of the restatement is checked against something independent of
it: closed forms, a brute-force truss on tiny graphs, networkx's k_truss
(goldens written in the build container), the reference's own kc_cpu binary
(goldens), textbook |N(u) ∩ N(v)| supports, bucket core numbers, a scipy
sparse-matrix truss peel, and a parser written here from the input contract.
"""
import json

import numpy as np
import pytest

from conftest import GOLDEN, canon_pairs
from graphs import brute_kmax_truss, closed_forms, random_graph
from oracle import oracle as O


def adjacency_sets(pairs, n):
    nb = [set() for _ in range(n)]
    for a, b in pairs:
        nb[int(a)].add(int(b))
        nb[int(b)].add(int(a))
    return nb


@pytest.mark.parametrize("case", closed_forms(), ids=lambda c: c[0])
def test_closed_forms(case):
    name, pairs, n, kmax, cnt = case
    s = O.solve_pairs(pairs, n, want_edges=True)
    assert (s.kmax, s.truss_edges) == (kmax, cnt)


@pytest.mark.parametrize("seed", range(40))
def test_brute_force_tiny(seed):
    n = 6 + seed % 14
    pairs = random_graph(n, 0.25 + 0.6 * ((seed * 3) % 7) / 7, 500 + seed)
    k, t = brute_kmax_truss(pairs, n)
    s = O.solve_pairs(pairs, n, want_edges=True)
    assert (s.kmax, s.truss_edges) == (k, len(t))
    np.testing.assert_array_equal(canon_pairs(s.truss), np.array(t, dtype=np.uint32).reshape(-1, 2))


def test_networkx_goldens():
    cases = json.loads((GOLDEN / "networkx_truss.json").read_text())
    assert len(cases) >= 30
    for c in cases:
        pairs = np.array(c["pairs"], dtype=np.uint32).reshape(-1, 2)
        s = O.solve_pairs(pairs, c["n"], want_edges=True)
        assert s.kmax == c["kmax"], c["seed"]
        want = np.array(c["truss"], dtype=np.uint32).reshape(-1, 2)
        np.testing.assert_array_equal(canon_pairs(s.truss), want)


KC_CPU_SMALL = ["er16k_clique24", "er3k_clique17", "rmat12", "rmat14_clique12_s7", "rmat16"]


@pytest.mark.parametrize("name", KC_CPU_SMALL)
def test_kc_cpu_goldens(name):
    from graphs import make_workload as make
    g = json.loads((GOLDEN / "kc_cpu_results.json").read_text())[name]
    pairs, n = make(g["params"])
    assert len(pairs) == g["m"]
    s = O.solve_pairs(pairs, n)
    assert (s.kmax, s.truss_edges) == (g["kmax"], g["truss_edges"])


def test_tiny_tsv_goldens():
    exp = json.loads((GOLDEN / "tiny" / "expected.json").read_text())
    assert exp
    for fname, line in exp.items():
        xs, ys = O.parse_tsv_bytes((GOLDEN / "tiny" / fname).read_bytes())
        off, adj, n = O.csr_from_lines(xs, ys)
        s = O.solve_csr(off, adj, n)
        assert f"kmax = {s.kmax}, Edges in kmax-truss = {s.truss_edges}." == line, fname


def _scipy_truss(pairs, n):
    """Independent truss peel: support = (A @ A) ∘ A with scipy sparse."""
    import scipy.sparse as sp
    p = np.asarray(pairs, dtype=np.int64)
    best_k, best = 0, None
    k = 3
    cur = p
    while len(cur):
        while True:
            A = sp.coo_matrix((np.ones(len(cur)), (cur[:, 0], cur[:, 1])), shape=(n, n)).tocsr()
            A = A + A.T
            S = (A @ A).multiply(A).tocsr()
            sup = np.asarray(S[cur[:, 0], cur[:, 1]]).ravel()
            keep = sup >= k - 2
            if keep.all():
                break
            cur = cur[keep]
            if not len(cur):
                break
        if not len(cur):
            break
        best_k, best = k, cur.copy()
        k += 1
    return best_k, best


def test_scipy_peel_pins_chung_lu_case():
    """The reference's kc_cpu never finishes on this graph (recorded in the
    golden file); the oracle is pinned here by an independent peel."""
    from graphs import make_workload as make
    g = json.loads((GOLDEN / "kc_cpu_results.json").read_text())["chung_lu_200k"]
    pairs, n = make(g["params"])
    s = O.solve_pairs(pairs, n, want_edges=True)
    assert (s.kmax, s.truss_edges) == (g.get("oracle_kmax", g.get("kmax")),
                                       g.get("oracle_truss_edges", g.get("truss_edges")))
    # independent check on the oracle's own (L-1)-core, which holds the truss
    off, adj, n = O.csr_from_pairs(pairs, n)
    removed = O.kcore_removed(off, adj, n, s.core_k)
    keep = (removed[pairs[:, 0]] == 0) & (removed[pairs[:, 1]] == 0)
    k, t = _scipy_truss(pairs[keep], n)
    assert (k, len(t)) == (s.kmax, s.truss_edges)
    np.testing.assert_array_equal(canon_pairs(t), canon_pairs(s.truss))


@pytest.mark.parametrize("seed", range(6))
def test_scipy_peel_random_medium(seed):
    pairs = random_graph(120, 0.15 + 0.05 * seed, 70 + seed, cliques=(6 + seed,))
    k, t = _scipy_truss(pairs, 120)
    s = O.solve_pairs(pairs, 120, want_edges=True)
    assert (s.kmax, s.truss_edges) == (k, len(t))
    np.testing.assert_array_equal(canon_pairs(s.truss), canon_pairs(t))


# ------------------------------------------------------ intermediate stages


def _graph(seed=3, n=300):
    return random_graph(n, 0.06, seed, cliques=(9, 7)), n


def test_support_matches_textbook():
    pairs, n = _graph()
    off, adj, n = O.csr_from_pairs(pairs, n)
    removed = np.zeros(n, dtype=np.uint8)
    nc, to, inv, degc = O.rank(off, adj, n, removed)
    off_d, adj_d = O.oriented(off, adj, n, to, nc)
    sup2 = O.support2(off_d, adj_d, nc)
    nb = adjacency_sets(pairs, n)
    for r in range(nc):
        for j in range(int(off_d[r]), int(off_d[r + 1])):
            u, v = int(inv[r]), int(inv[adj_d[j]])
            assert sup2[j] == 2 * len(nb[u] & nb[v])


def test_core_matches_bucket_core_numbers():
    for seed in range(5):
        pairs, n = _graph(seed, 400)
        off, adj, n = O.csr_from_pairs(pairs, n)
        core = O.core_numbers(off, adj, n)
        for k in (1, 2, 3, 5, 8, int(core.max())):
            removed = O.kcore_removed(off, adj, n, k)
            np.testing.assert_array_equal(removed == 0, core >= k)
        kmax, removed = O.max_core(off, adj, n)
        assert kmax == core.max()
        np.testing.assert_array_equal(removed == 0, core == kmax)


def test_core_numbers_match_networkx():
    nx = pytest.importorskip("networkx")
    pairs, n = _graph(11, 500)
    off, adj, n = O.csr_from_pairs(pairs, n)
    core = O.core_numbers(off, adj, n)
    g = nx.Graph()
    g.add_nodes_from(range(n))
    g.add_edges_from(map(tuple, pairs.tolist()))
    ref = nx.core_number(g)
    assert [ref[v] for v in range(n)] == core.tolist()


def test_rank_orientation_and_rows():
    pairs, n = _graph(5, 250)
    off, adj, n = O.csr_from_pairs(pairs, n)
    removed = O.kcore_removed(off, adj, n, 3)
    nc, to, inv, degc = O.rank(off, adj, n, removed)
    kept = np.nonzero(removed == 0)[0]
    # rank order = (degree inside the kept set ascending, id descending)
    order = kept[np.lexsort((-kept.astype(np.int64), degc[kept]))]
    np.testing.assert_array_equal(inv, order)
    off_d, adj_d = O.oriented(off, adj, n, to, nc)
    nb = adjacency_sets(pairs, n)
    for r in range(nc):
        row = adj_d[off_d[r]:off_d[r + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)
        want = sorted(int(to[w]) for w in nb[int(inv[r])] if to[w] != 0xFFFFFFFF and to[w] > r)
        assert row.tolist() == want


def test_undirected_view():
    pairs, n = _graph(8, 200)
    off, adj, n = O.csr_from_pairs(pairs, n)
    nc, to, inv, degc = O.rank(off, adj, n, np.zeros(n, dtype=np.uint8))
    off_d, adj_d = O.oriented(off, adj, n, to, nc)
    off_u, adj_u, fa, lson, rson = O.undirected(off_d, adj_d, nc)
    src = np.repeat(np.arange(nc), np.diff(off_d).astype(np.int64))
    for w in range(nc):
        row = adj_u[off_u[w]:off_u[w + 1]].astype(np.int64)
        assert np.all(np.diff(row) > 0)
    for e in range(len(adj_d)):
        u, v = src[e], adj_d[e]
        assert off_u[u] <= lson[e] < off_u[u + 1] and adj_u[lson[e]] == v and fa[lson[e]] == e
        assert off_u[v] <= rson[e] < off_u[v + 1] and adj_u[rson[e]] == u and fa[rson[e]] == e


def test_parse_matches_contract():
    text = b"1\t0\n2 0\r\n0\t1 extra cols\n2\t1\n0,2\n1\t2\n"
    xs, ys = O.parse_tsv_bytes(text)
    assert xs.tolist() == [1, 2, 0, 2, 0, 1]
    assert ys.tolist() == [0, 0, 1, 1, 2, 2]
    off, adj, n = O.csr_from_lines(xs, ys)
    assert n == 3 and off.tolist() == [0, 2, 4, 6]
    s = O.solve_csr(off, adj, n)
    assert (s.kmax, s.truss_edges) == (3, 3)


def test_planted_support_error_is_caught():
    """A dropped term in the support count must fail the textbook pin."""
    pairs, n = _graph(2, 150)
    off, adj, n = O.csr_from_pairs(pairs, n)
    nc, to, inv, _ = O.rank(off, adj, n, np.zeros(n, dtype=np.uint8))
    off_d, adj_d = O.oriented(off, adj, n, to, nc)
    sup2 = O.support2(off_d, adj_d, nc)
    nb = adjacency_sets(pairs, n)
    j = int(np.argmax(sup2))
    bad = sup2.copy()
    bad[j] -= 2
    r = int(np.searchsorted(off_d, j, side="right") - 1)
    u, v = int(inv[r]), int(inv[adj_d[j]])
    assert bad[j] != 2 * len(nb[u] & nb[v])


def _truss_by_definition(edges):
    """t(e) = the largest k whose k-truss holds e, each k-truss found as the
    fixpoint of deleting edges with fewer than k - 2 triangles (supports
    recounted from scratch every pass)."""
    out = {e: 2 for e in edges}
    k = 3
    cur = set(edges)
    while cur:
        while True:
            adj = {}
            for u, v in cur:
                adj.setdefault(u, set()).add(v)
                adj.setdefault(v, set()).add(u)
            drop = {(u, v) for u, v in cur if len(adj[u] & adj[v]) < k - 2}
            if not drop:
                break
            cur -= drop
        for e in cur:
            out[e] = k
        k += 1
    return out


def test_trussness_matches_definition_on_small_graphs():
    import itertools
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(4, 13))
        p = float(rng.uniform(0.2, 0.9))
        edges = [(a, b) for a, b in itertools.combinations(range(n), 2) if rng.random() < p]
        if not edges:
            continue
        got = O.trussness(np.array(edges))
        assert got == _truss_by_definition(edges), trial


def test_trussness_closed_forms():
    import itertools
    def clique(vs):
        return [(a, b) for a, b in itertools.combinations(vs, 2)]
    k5_tail = clique(range(5)) + [(4, 5), (5, 6)]
    t = O.trussness(np.array(k5_tail))
    assert all(t[e] == 5 for e in clique(range(5))) and t[(4, 5)] == 2 and t[(5, 6)] == 2
    parts = clique(range(4)) + clique(range(10, 16)) + clique(range(20, 23))
    t = O.trussness(np.array(parts))
    assert all(t[e] == 4 for e in clique(range(4)))
    assert all(t[e] == 6 for e in clique(range(10, 16)))
    assert all(t[e] == 3 for e in clique(range(20, 23)))
    cycle = [(i, (i + 1) % 9) for i in range(9)]
    assert set(O.trussness(np.array(cycle)).values()) == {2}


def test_trussness_top_level_is_the_kmax_truss():
    """The edges of the largest trussness are the oracle's kmax-truss (pinned
    against kc_cpu and networkx elsewhere), on random graphs with cliques."""
    for seed in range(6):
        pairs, n = _graph(seed, 300)
        t = O.trussness(pairs)
        s = O.solve_pairs(pairs, n, want_edges=True)
        top = max(t.values())
        kmax = top if top >= 3 else 0
        assert kmax == s.kmax
        got = sorted(e for e, k in t.items() if k == top) if kmax else []
        want = sorted(map(tuple, np.sort(s.truss, axis=1).tolist())) if kmax else []
        assert got == want
