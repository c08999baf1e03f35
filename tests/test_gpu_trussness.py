"""Edge-set outputs: per-edge trussness from the device decomposition equals
the oracle's textbook peel (itself pinned against the definition of the
k-truss), its top level is the kmax-truss of the plain solve, and the CLI's
-o / -t files hold exactly those edges."""
import subprocess

import numpy as np
import pytest

from conftest import canon_pairs
from graphs import closed_forms, random_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle_rows(pairs) -> np.ndarray:
    t = O.trussness(pairs)
    rows = sorted((u, v, k) for (u, v), k in t.items() if k >= 3)
    return np.array(rows, dtype=np.uint32).reshape(-1, 3)


def _cases():
    out = [(name, np.asarray(p, np.uint32).reshape(-1, 2), n) for name, p, n, _, _ in closed_forms()]
    for seed in range(4):
        out.append((f"er{seed}", random_graph(400, 0.05, seed, cliques=(11, 8)), 400))
    from kc_cuda_amd import gen
    for seed in (1, 2):
        p, n = gen.rmat(11, 8, 0.57, 0.19, 0.19, True, 9, seed)
        out.append((f"rmat11_{seed}", p, n))
    return out


def test_trussness_matches_oracle(ctx):
    for name, pairs, n in _cases():
        print(f"[gpu] trussness {name} m={len(pairs)}", flush=True)
        r = ctx.solve_pairs(pairs, n)
        t = ctx.trussness()
        np.testing.assert_array_equal(t, _oracle_rows(pairs), err_msg=name)
        if r.kmax:
            assert int(t[:, 2].max()) == r.kmax and int((t[:, 2] == r.kmax).sum()) == r.truss_edges
        else:
            assert len(t) == 0


def test_trussness_top_level_on_rmat16(ctx):
    """At a size the Python oracle does not reach: the top level equals the
    plain solve's survivors, and every listed edge is a real edge."""
    from kc_cuda_amd import gen
    pairs, n = gen.rmat(16, 16, 0.57, 0.19, 0.19, True, 20, 4)
    r = ctx.solve_pairs(pairs, n)
    surv = ctx.truss_edges()
    t = ctx.trussness()
    top = t[t[:, 2] == r.kmax][:, :2]
    np.testing.assert_array_equal(top, surv)
    # the decomposition does not take the solve's kmax-truss away, and a
    # second export copies the same triples
    np.testing.assert_array_equal(ctx.truss_edges(), surv)
    np.testing.assert_array_equal(ctx.trussness(), t)
    assert int(t[:, 2].max()) == r.kmax
    canon = canon_pairs(pairs)
    key = canon[:, 0].astype(np.uint64) << np.uint64(32) | canon[:, 1]
    tk = t[:, 0].astype(np.uint64) << np.uint64(32) | t[:, 1]
    assert np.isin(tk, key).all() and len(np.unique(tk)) == len(tk)


def test_cli_edge_and_trussness_files(tmp_path):
    from kc_cuda_amd import gen
    from kc_cuda_amd._lib import kc_gpu_binary
    pairs = random_graph(300, 0.06, 5, cliques=(10, 7))
    path = tmp_path / "g.tsv"
    gen.write_tsv(path, pairs)
    eo, to = tmp_path / "edges.tsv", tmp_path / "truss.tsv"
    out = subprocess.run([str(kc_gpu_binary()), "-f", str(path), "-o", str(eo), "-t", str(to)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    s = O.solve_pairs(pairs, 300, want_edges=True)
    assert out.stdout.strip().splitlines()[-1] == \
        f"kmax = {s.kmax}, Edges in kmax-truss = {s.truss_edges}."
    edges = np.loadtxt(eo, dtype=np.uint32, ndmin=2).reshape(-1, 2)
    np.testing.assert_array_equal(edges, canon_pairs(s.truss))
    truss = np.loadtxt(to, dtype=np.uint32, ndmin=2).reshape(-1, 3)
    np.testing.assert_array_equal(truss, _oracle_rows(pairs))
    bad = subprocess.run([str(kc_gpu_binary()), "-f", str(path), "-x", "y"], capture_output=True, text=True,
                         timeout=60)
    assert bad.returncode == 1
