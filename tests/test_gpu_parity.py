"""GPU parity: the HIP path through the C ABI against the oracle and the
committed golden fixtures. Results must be bit-exact: same kmax, same truss
size and the same surviving edge set."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, canon_pairs
from graphs import closed_forms, random_graph

pytestmark = pytest.mark.gpu


def _oracle():
    from oracle import oracle as O
    return O


@pytest.mark.parametrize("case", closed_forms(), ids=lambda c: c[0])
def test_closed_forms(ctx, case):
    name, pairs, n, kmax, cnt = case
    print(f"[gpu] closed form {name} n={n} m={len(pairs)}", flush=True)
    r = ctx.solve_pairs(pairs, n)
    assert (r.kmax, r.truss_edges) == (kmax, cnt)
    t = ctx.truss_edges()
    assert t.shape[0] == cnt


@pytest.mark.parametrize("seed", range(12))
def test_random_small_vs_oracle(ctx, seed):
    O = _oracle()
    n = [20, 40, 80, 150, 300, 600][seed % 6]
    p = [0.3, 0.2, 0.1, 0.05, 0.03, 0.02][seed % 6]
    pairs = random_graph(n, p, seed, cliques=(seed % 5 + 3,))
    print(f"[gpu] random seed={seed} n={n} m={len(pairs)}", flush=True)
    ref = O.solve_pairs(pairs, n, want_edges=True)
    r = ctx.solve_pairs(pairs, n)
    assert (r.kmax, r.truss_edges) == (ref.kmax, ref.truss_edges)
    assert r.max_core == ref.max_core
    assert r.k_hat == ref.k_hat
    np.testing.assert_array_equal(ctx.truss_edges(), canon_pairs(ref.truss))


def test_config1_er_clique(ctx):
    from kc_cuda_amd import gen
    pairs, n = gen.er_clique(16384, 16.0, 24, 0)
    print(f"[gpu] ER(16384,16)+K24 m={len(pairs)}", flush=True)
    gold = json.loads((GOLDEN / "kc_cpu_results.json").read_text())["er16k_clique24"]
    assert len(pairs) == gold["m"]
    r = ctx.solve_pairs(pairs, n)
    assert (r.kmax, r.truss_edges) == (gold["kmax"], gold["truss_edges"])
    ref = _oracle().solve_pairs(pairs, n, want_edges=True)
    np.testing.assert_array_equal(ctx.truss_edges(), canon_pairs(ref.truss))


@pytest.mark.parametrize("lanes", [8, 16, 32, 64])
@pytest.mark.parametrize("force_edge", [False, True])
def test_kernel_variants_agree(ctx, lanes, force_edge):
    from kc_cuda_amd import gen
    pairs, n = gen.rmat(14, 16, 0.57, 0.19, 0.19, True, 12, 7)
    print(f"[gpu] rmat14 lanes={lanes} edge={force_edge}", flush=True)
    ref = _oracle().solve_pairs(pairs, n, want_edges=True)
    try:
        ctx.tune(lanes, lanes, force_edge)
        r = ctx.solve_pairs(pairs, n)
        assert (r.kmax, r.truss_edges) == (ref.kmax, ref.truss_edges)
        np.testing.assert_array_equal(ctx.truss_edges(), canon_pairs(ref.truss))
    finally:
        ctx.tune(0, 0, False)


@pytest.mark.parametrize("core", ["sparse", "dense", "dense-walk", "dense-probe"])
@pytest.mark.parametrize("seed", range(8))
def test_core_representations_agree(ctx, core, seed):
    """Sorted-row and bit-matrix cores give the same kmax-truss, and on
    bit-matrix cores the first pass's threshold probes give the same k-hat as
    walking every level."""
    from kc_cuda_amd import gen
    O = _oracle()
    if seed < 4:
        pairs = random_graph([60, 200, 500, 900][seed], [0.3, 0.1, 0.04, 0.02][seed], 40 + seed,
                             cliques=(seed + 5, 4))
        n = [60, 200, 500, 900][seed]
    else:
        pairs, n = gen.rmat(10 + seed, 16, 0.57, 0.19, 0.19, True, 6 * seed, 90 + seed)
    print(f"[gpu] core={core} seed={seed} m={len(pairs)}", flush=True)
    ref = O.solve_pairs(pairs, n, want_edges=True)
    try:
        core_kind, _, pass1 = core.partition("-")
        ctx.tune(core=core_kind, pass1=pass1 or "auto")
        r = ctx.solve_pairs(pairs, n)
        assert (r.kmax, r.truss_edges) == (ref.kmax, ref.truss_edges)
        assert (r.k_hat, r.k_hat_edges) == (ref.k_hat, ref.k_hat_edges)
        np.testing.assert_array_equal(ctx.truss_edges(), canon_pairs(ref.truss))
    finally:
        ctx.tune()


def _vertex_cases():
    """Graphs for the vertex-peel variants: random with cliques, a star whose
    hub row spans many cascade chunks, two stars sharing leaves, R-MAT with
    hub rows above 4096 slots, and ER with a planted clique."""
    from kc_cuda_amd import gen
    star = np.array([[0, v] for v in range(1, 20000)] + [[1, 2], [2, 3], [1, 3]], dtype=np.uint32)
    two = np.array([[0, v] for v in range(2, 9000)] + [[1, v] for v in range(2, 9000)] + [[0, 1]],
                   dtype=np.uint32)
    return [("random", random_graph(700, 0.03, 77, cliques=(9, 6)), 700),
            ("star", canon_pairs(star), 20000),
            ("two_stars", canon_pairs(two), 9000),
            ("rmat17", *gen.rmat(17, 16, 0.57, 0.19, 0.19, True, 20, 13)),
            ("er_clique", *gen.er_clique(30000, 12.0, 30, 9))]


@pytest.mark.parametrize("vertex", ["cascade", "rounds", "auto"])
def test_vertex_peel_variants_agree(ctx, vertex):
    """The vertex peel as one cascade launch per threshold and as
    synchronous rounds: the same max core, k-hat, kmax and kmax-truss as the
    oracle."""
    O = _oracle()
    for name, pairs, n in _vertex_cases():
        print(f"[gpu] vertex={vertex} {name} n={n} m={len(pairs)}", flush=True)
        ref = O.solve_pairs(pairs, n, want_edges=True)
        try:
            ctx.tune(vertex=vertex)
            r = ctx.solve_pairs(pairs, n)
            edges = ctx.truss_edges()
        finally:
            ctx.tune()
        assert (r.max_core, r.k_hat, r.kmax, r.truss_edges) == (ref.max_core, ref.k_hat, ref.kmax,
                                                                ref.truss_edges), name
        np.testing.assert_array_equal(edges, canon_pairs(ref.truss), err_msg=name)


@pytest.mark.parametrize("scale", [12, 16, 18, 20])
def test_rmat_vs_golden(ctx, scale):
    from kc_cuda_amd import gen
    gold = json.loads((GOLDEN / "kc_cpu_results.json").read_text())[f"rmat{scale}"]
    pairs, n = gen.rmat(scale)
    print(f"[gpu] rmat{scale} m={len(pairs)}", flush=True)
    assert len(pairs) == gold["m"]
    r = ctx.solve_pairs(pairs, n)
    assert (r.kmax, r.truss_edges) == (gold["kmax"], gold["truss_edges"])
    if scale <= 18:
        ref = _oracle().solve_pairs(pairs, n, want_edges=True)
        np.testing.assert_array_equal(ctx.truss_edges(), canon_pairs(ref.truss))


def test_compaction_runs_on_rmat16(ctx):
    """The edge peel compacts its edge set (src/kc.cu:1059-1138) on a sparse
    core: forced sorted-row cores of R-MAT 16 rewrite the edge set at least
    once, and the answer still equals the oracle's."""
    from kc_cuda_amd import gen
    pairs, n = gen.rmat(16)
    ref = _oracle().solve_pairs(pairs, n, want_edges=True)
    try:
        ctx.tune(core="sparse")
        r = ctx.solve_pairs(pairs, n)
    finally:
        ctx.tune()
    print(f"[gpu] rmat16 sparse: compactions={r.compactions} recounts={r.recounts}", flush=True)
    assert r.compactions > 0
    assert (r.kmax, r.truss_edges) == (ref.kmax, ref.truss_edges)
    np.testing.assert_array_equal(ctx.truss_edges(), canon_pairs(ref.truss))


@pytest.mark.parametrize("atomic", [False, True])
def test_filter_cut_paths(capfd, atomic):
    """The degree filter's cut of R-MAT 16 (h-index 656: the vertices of
    degree >= 41 are kept, their alive neighbours counted from their rows).
    After the bucketed build its partition buffer holds every kept row, so
    the rows are filled at full length while being counted, then packed (one
    walk); on a fresh context with the atomic build there is no such buffer
    and the cut counts, then fills (two walks). Both equal the oracle."""
    from kc_cuda_amd import Context, gen
    pairs, n = gen.rmat(16)
    ref = _oracle().solve_pairs(pairs, n, want_edges=True)
    old = os.environ.get("KT_TRACE")
    os.environ["KT_TRACE"] = "1"  # read when the context is made
    try:
        c = Context(0)
    finally:
        if old is None:
            os.environ.pop("KT_TRACE")
        else:
            os.environ["KT_TRACE"] = old
    try:
        if atomic:
            c.tune(force_atomic_build=True)
        capfd.readouterr()
        r = c.solve_pairs(pairs, n)
        edges = c.truss_edges()
        err = capfd.readouterr().err
    finally:
        c.close()
    cuts = [l for l in err.splitlines() if l.startswith("[kt] cut ") and "recount=1" in l]
    print(f"[gpu] filter cut atomic={atomic}: {cuts}", flush=True)
    assert cuts, "the degree filter did not cut"
    want = "fill=count+fill" if atomic else "fill=one-walk"
    assert all(want in l for l in cuts)
    assert (r.kmax, r.truss_edges) == (ref.kmax, ref.truss_edges)
    np.testing.assert_array_equal(edges, canon_pairs(ref.truss))


@pytest.mark.parametrize("build", ["auto", "atomic", "unsorted"])
@pytest.mark.parametrize("order", ["sorted", "shuffled", "flipped"])
@pytest.mark.parametrize("lines", [False, True])
def test_csr_matches_oracle(ctx, build, order, lines):
    """CSR rows (as sets) and offsets against the oracle's, for inputs in the
    reference's order (lines sorted by row, pairs sorted: the builds that
    read rows as runs), shuffled (the bucketed build), and, for pairs, with
    a random half flipped to (v, u) and then sorted by first id (sorted by
    first id but not canonical)."""
    from kc_cuda_amd import gen
    O = _oracle()
    pairs, n = gen.rmat(12, 8, 0.57, 0.19, 0.19, True, 0, 5)
    rng = np.random.default_rng(9)
    if order == "flipped":
        if lines:
            pytest.skip("lines hold both directions already")
        flip = rng.random(len(pairs)) < 0.5
        pairs = pairs.copy()
        pairs[flip] = pairs[flip][:, ::-1]
        pairs = pairs[np.argsort(pairs[:, 0], kind="stable")]
    print(f"[gpu] csr build={build} order={order} lines={lines}", flush=True)
    try:
        ctx.tune(force_atomic_build=build == "atomic", unsorted_build=build == "unsorted")
        if lines:
            xs, ys = O.lines_from_pairs(pairs)
            if order == "shuffled":
                p = rng.permutation(len(xs))
                xs, ys = np.ascontiguousarray(xs[p]), np.ascontiguousarray(ys[p])
            ctx.solve_lines(xs, ys, n)
        else:
            q = pairs[rng.permutation(len(pairs))] if order == "shuffled" else pairs
            ctx.solve_pairs(np.ascontiguousarray(q), n)
        off, adj = ctx.last_csr(n, 2 * len(pairs))
    finally:
        ctx.tune()
    ooff, oadj, _ = O.csr_from_pairs(canon_pairs(pairs), n)
    np.testing.assert_array_equal(off, ooff)
    for v in range(n):  # rows are sets: the device scatter order is free
        a = np.sort(adj[off[v]:off[v + 1]])
        b = np.sort(oadj[ooff[v]:ooff[v + 1]])
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("lines", [False, True])
def test_sorted_build_multi_bucket(ctx, lines):
    """Row-sorted inputs over many buckets (n = 2^18: 256 buckets of 1024
    rows, with empty rows at both ends and inside) give the bucketed build's
    CSR, and the same solve."""
    from kc_cuda_amd import gen
    O = _oracle()
    pairs, _ = gen.rmat(16, 8, 0.57, 0.19, 0.19, True, 0, 12)
    n = 1 << 18
    pairs = canon_pairs((pairs.astype(np.uint64) * 3 + 7).astype(np.uint32))  # ids spread out, gaps
    print(f"[gpu] sorted multi-bucket lines={lines} m={len(pairs)}", flush=True)
    got = {}
    for build in ("auto", "unsorted"):
        try:
            ctx.tune(unsorted_build=build == "unsorted")
            if lines:
                xs, ys = O.lines_from_pairs(pairs)
                r = ctx.solve_lines(xs, ys, n)
            else:
                r = ctx.solve_pairs(pairs, n)
            off, adj = ctx.last_csr(n, 2 * len(pairs))
        finally:
            ctx.tune()
        rows = np.repeat(np.arange(n, dtype=np.uint64), np.diff(off.astype(np.int64)))
        key = np.sort((rows << np.uint64(32)) | adj.astype(np.uint64))
        got[build] = (off, key, r.kmax, r.truss_edges)
    np.testing.assert_array_equal(got["auto"][0], got["unsorted"][0])
    np.testing.assert_array_equal(got["auto"][1], got["unsorted"][1])
    assert got["auto"][2:] == got["unsorted"][2:]
    ooff, _, _ = O.csr_from_pairs(pairs, n)
    np.testing.assert_array_equal(got["auto"][0], ooff)


def test_tsv_path_and_cli(ctx, tmp_path):
    from kc_cuda_amd import gen, parse_tsv
    from kc_cuda_amd._lib import kc_gpu_binary
    O = _oracle()
    pairs, n = gen.er_clique(3000, 10.0, 17, 11)
    path = tmp_path / "g.tsv"
    gen.write_tsv(path, pairs)
    xs, ys, nn = parse_tsv(path)
    oxs, oys = O.parse_tsv_bytes(path.read_bytes())
    np.testing.assert_array_equal(xs, oxs)
    np.testing.assert_array_equal(ys, oys)
    ref = O.solve_pairs(pairs, n)
    r, _ = ctx.solve_tsv(path)
    assert (r.kmax, r.truss_edges) == (ref.kmax, ref.truss_edges)
    out = subprocess.run([str(kc_gpu_binary()), "-f", str(path)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == \
        f"kmax = {ref.kmax}, Edges in kmax-truss = {ref.truss_edges}."


def test_repeat_determinism_and_relabel(ctx):
    from kc_cuda_amd import gen
    pairs, n = gen.rmat(15, 16, 0.57, 0.19, 0.19, True, 20, 3)
    r1 = ctx.solve_pairs(pairs, n)
    t1 = ctx.truss_edges()
    r2 = ctx.solve_pairs(pairs, n)
    t2 = ctx.truss_edges()
    assert (r1.kmax, r1.truss_edges) == (r2.kmax, r2.truss_edges)
    np.testing.assert_array_equal(t1, t2)
    perm = np.random.default_rng(9).permutation(n).astype(np.uint32)
    r3 = ctx.solve_pairs(canon_pairs(perm[pairs]), n)
    assert (r3.kmax, r3.truss_edges) == (r1.kmax, r1.truss_edges)
    inv = np.argsort(perm).astype(np.uint32)
    np.testing.assert_array_equal(canon_pairs(inv[ctx.truss_edges()]), t1)


def test_device_resident_input(ctx):
    from kc_cuda_amd import gen
    pairs, n = gen.rmat(13)
    d = ctx.to_device(np.ascontiguousarray(pairs))
    try:
        r = ctx.solve_device(d, len(pairs), n)
    finally:
        d.free()
    r2 = ctx.solve_pairs(pairs, n)
    assert (r.kmax, r.truss_edges) == (r2.kmax, r2.truss_edges)


def test_bad_vertex_id_raises(ctx):
    from kc_cuda_amd import KtError
    pairs = np.array([[0, 1], [1, 7]], dtype=np.uint32)
    with pytest.raises(KtError):
        ctx.solve_pairs(pairs, 5)


def test_push_probe_keeps_results(ctx):
    """The measurement hook of the vertex-peel push (events at each launch's
    start and end, slot counters) leaves every answer unchanged."""
    from kc_cuda_amd import gen
    pairs, n = gen.rmat(15, 16, 0.57, 0.19, 0.19, True, 16, 8)
    ref = ctx.solve_pairs(pairs, n)
    t_ref = ctx.truss_edges()
    try:
        ctx.push_probe(1)
        r = ctx.solve_pairs(pairs, n)
        timed = ctx.push_stats()
        np.testing.assert_array_equal(ctx.truss_edges(), t_ref)
        ctx.push_probe(3)
        r2 = ctx.solve_pairs(pairs, n)
        both = ctx.push_stats()
    finally:
        ctx.push_probe(0)
    for x in (r, r2):
        assert (x.kmax, x.truss_edges, x.max_core, x.k_hat) == (ref.kmax, ref.truss_edges, ref.max_core, ref.k_hat)
    np.testing.assert_array_equal(ctx.truss_edges(), t_ref)
    assert timed["launches"] > 0 and timed["ms"] > 0 and timed["slots"] == 0
    assert both["launches"] == timed["launches"] and 0 < both["slots"] <= 2 * len(pairs) * 4
    assert 0 < both["vertices"]
