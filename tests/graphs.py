"""Small hand-built graphs with closed-form k-max-truss answers, and tiny
random graphs, shared by the CPU and GPU test suites."""
from __future__ import annotations

import itertools

import numpy as np


def clique(vs) -> list[tuple[int, int]]:
    vs = list(vs)
    return [(min(a, b), max(a, b)) for a, b in itertools.combinations(vs, 2)]


def as_pairs(edges) -> np.ndarray:
    s = sorted({(min(a, b), max(a, b)) for a, b in edges if a != b})
    return np.array(s, dtype=np.uint32).reshape(-1, 2)


def closed_forms():
    """(name, pairs, n, kmax, truss_edges); kmax 0 means triangle-free."""
    out = []
    for k in (3, 4, 5, 8, 13, 24):
        out.append((f"K{k}", as_pairs(clique(range(k))), k, k, k * (k - 1) // 2))
    # disjoint cliques: the largest wins
    e = clique(range(4)) + clique(range(10, 16)) + clique(range(20, 23))
    out.append(("K4+K6+K3", as_pairs(e), 23, 6, 15))
    # K5 with a tail
    e = clique(range(5)) + [(4, 5), (5, 6), (6, 7)]
    out.append(("K5+tail", as_pairs(e), 8, 5, 10))
    # triangle-free: path, cycle, star, bipartite
    out.append(("path", as_pairs([(i, i + 1) for i in range(9)]), 10, 0, 0))
    out.append(("cycle", as_pairs([(i, (i + 1) % 12) for i in range(12)]), 12, 0, 0))
    out.append(("star", as_pairs([(0, i) for i in range(1, 30)]), 30, 0, 0))
    out.append(("K20,20", as_pairs([(a, 20 + b) for a in range(20) for b in range(20)]), 40, 0, 0))
    out.append(("edge", as_pairs([(0, 1)]), 2, 0, 0))
    out.append(("empty", as_pairs([]), 5, 0, 0))
    # a dense triangle-free block (its max-core) plus a low-core triangle
    e = [(a, 20 + b) for a in range(20) for b in range(20)] + [(50, 51), (51, 52), (50, 52)]
    out.append(("K20,20+triangle", as_pairs(e), 53, 3, 3))
    # two cliques sharing an edge: K6 and K6 on {0,1} + 4 others each
    e = clique([0, 1, 2, 3, 4, 5]) + clique([0, 1, 6, 7, 8, 9])
    out.append(("K6-K6 shared edge", as_pairs(e), 10, 6, 29))
    # wheel W_8 (hub + 8-cycle): every edge in a triangle, no 4-truss
    e = [(0, i) for i in range(1, 9)] + [(i, i % 8 + 1) for i in range(1, 9)]
    out.append(("wheel8", as_pairs(e), 9, 3, 16))
    # isolated high ids
    out.append(("K4 high ids", as_pairs(clique([100, 200, 300, 999]))
                , 1000, 4, 6))
    return out


def random_graph(n: int, p: float, seed: int, cliques=()) -> np.ndarray:
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    keep = rng.random(iu[0].size) < p
    edges = list(zip(iu[0][keep].tolist(), iu[1][keep].tolist()))
    for size in cliques:
        vs = rng.choice(n, size=min(size, n), replace=False)
        edges += clique(vs.tolist())
    return as_pairs(edges)


def brute_kmax_truss(pairs: np.ndarray, n: int):
    """Textbook definition: for k = 3, 4, ..., repeatedly drop edges in fewer
    than k-2 triangles of the remaining graph; kmax is the last k with edges
    left. Returns (kmax, sorted edge list of the kmax-truss)."""
    edges = {(int(a), int(b)) for a, b in pairs}
    best_k, best = 0, []
    k = 3
    cur = set(edges)
    while True:
        changed = True
        while changed and cur:
            nbr = {}
            for a, b in cur:
                nbr.setdefault(a, set()).add(b)
                nbr.setdefault(b, set()).add(a)
            drop = {(a, b) for a, b in cur if len(nbr[a] & nbr[b]) < k - 2}
            changed = bool(drop)
            cur -= drop
        if not cur:
            break
        best_k, best = k, sorted(cur)
        k += 1
    return best_k, best


def make_workload(params: dict):
    """Generate a graph from a golden-fixture parameter record."""
    from kc_cuda_amd import gen
    p = dict(params)
    kind = p.pop("kind")
    if kind == "er_clique":
        return gen.er_clique(**p)
    if kind == "rmat":
        return gen.rmat(**p)
    if kind == "chung_lu":
        return gen.chung_lu(**p)
    raise KeyError(kind)
