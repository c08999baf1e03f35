"""Parity at the benchmark workloads' full sizes, against answers the oracle
computed on the same generated graphs (tests/golden/full_configs.json, made
by tests/golden/make_full_goldens.py) and, for RMAT-22, the reference's own
kc_cpu (tests/golden/kc_cpu_results.json):

  rmat22        plain solve, 64M edges: kmax, |T| and the kmax-truss edge set
  chung_lu_50m  plain solve, 500M edges (uint64 CSR offsets): the same
  batch4096     the 4096-graph mix in one batched solve: every graph's kmax,
                |T| and kmax-truss edge set (kept-pair flags)
  serve160      the 160-query serving mix through 8 slots: every query's
                kmax and |T| in the mix, alone, and against the oracle

Each test first checks that the generator reproduced the graphs the oracle
saw (sha256 of the pairs), so a mismatch names the solver, not the inputs."""
import hashlib
import json

import numpy as np
import pytest

from conftest import GOLDEN, canon_pairs

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

KC_CPU = json.loads((GOLDEN / "kc_cpu_results.json").read_text())


def _golden(name: str) -> dict:
    path = GOLDEN / "full_configs.json"
    assert path.exists(), f"{path} is missing: run tests/golden/make_full_goldens.py"
    data = json.loads(path.read_text())
    assert name in data, f"{name} is missing from {path}"
    return data[name]


def _sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u4").tobytes()).hexdigest()


def _say(msg: str) -> None:
    print(f"[gpu] {msg}", flush=True)


@pytest.mark.parametrize("name", ["rmat22", "chung_lu_50m"])
def test_plain_full_size(ctx, name):
    from kc_cuda_amd import gen
    want = _golden(name)
    pairs, n = gen.workload(name)
    assert (n, len(pairs)) == (want["n"], want["m"])
    assert _sha(pairs) == want["pairs_sha256"], "generator output changed"
    _say(f"{name}: n={n} m={len(pairs)} solve from device-resident pairs")
    d = ctx.to_device(pairs)
    del pairs
    try:
        r = ctx.solve_device(d, want["m"], n)
    finally:
        d.free()
    _say(f"{name}: kmax={r.kmax} |T|={r.truss_edges} {r.ms_total:.1f} ms")
    assert (r.kmax, r.truss_edges) == (want["kmax"], want["truss_edges"])
    if name in KC_CPU and KC_CPU[name].get("kmax") is not None:
        assert (r.kmax, r.truss_edges) == (KC_CPU[name]["kmax"], KC_CPU[name]["truss_edges"])
    edges = ctx.truss_edges()
    assert len(edges) == want["truss_edges"]
    assert _sha(edges) == want["truss_sha256"], "kmax-truss edge set differs from the oracle's"
    st = ctx.stats()
    assert st.n_core > 0 and st.m_core >= r.truss_edges


def test_batch4096_full_size(ctx):
    from kc_cuda_amd import gen
    want = _golden("batch4096")
    pairs, offs, ns = gen.batch()
    assert len(ns) == want["graphs"] == 4096
    np.testing.assert_array_equal(ns, want["n"])
    np.testing.assert_array_equal(np.diff(offs.astype(np.int64)), want["m"])
    assert _sha(pairs) == want["pairs_sha256"], "generator output changed"
    _say(f"batch4096: {len(ns)} graphs, {len(pairs)} edges, one batched solve")
    d = ctx.to_device(pairs)
    try:
        res = ctx.batch_solve(d, offs, ns, keep=True)
    finally:
        d.free()
    assert (res.status == 0).all()
    bad = np.nonzero((res.kmax != np.asarray(want["kmax"])) |
                     (res.edges != np.asarray(want["truss_edges"])))[0]
    assert bad.size == 0, f"graphs {bad[:10].tolist()} differ from the oracle"
    h = hashlib.sha256()
    for g in range(len(ns)):
        q = pairs[int(offs[g]):int(offs[g + 1])]
        kept = canon_pairs(q[res.keep[int(offs[g]):int(offs[g + 1])] == 1])
        assert len(kept) == want["truss_edges"][g], f"graph {g}"
        if len(kept):
            tagged = np.empty((len(kept), 3), dtype="<u4")
            tagged[:, 0] = g
            tagged[:, 1:] = kept
            h.update(tagged.tobytes())
    assert h.hexdigest() == want["truss_sha256"], "kmax-truss edge sets differ from the oracle's"


def test_serve160_full_size(ctx):
    """Each query's answer in the served mix equals its stand-alone solve and
    the oracle's; the stand-alone solves' edge sets equal the oracle's."""
    from kc_cuda_amd import Server, gen
    want = _golden("serve160")
    mix = gen.serve_mix(160, 200_000, 1.5)
    assert [n for _, n in mix] == want["n"]
    assert [len(p) for p, _ in mix] == want["m"]
    hp = hashlib.sha256()
    for p, _ in mix:
        hp.update(np.ascontiguousarray(p, dtype="<u4").tobytes())
    assert hp.hexdigest() == want["pairs_sha256"], "generator output changed"
    _say(f"serve160: {len(mix)} queries, {sum(want['m'])} edges, 8 slots")
    with Server(0, 8, 96 << 30) as srv:
        tickets = [srv.submit(p, n) for p, n in mix]
        served = [srv.wait(t) for t in tickets]
        c = srv.counters()
    assert c["completed"] == len(mix) and c["failed"] == 0
    _say(f"serve160: served, {c['grouped']} grouped; now alone")
    h = hashlib.sha256()
    for i, ((p, n), s) in enumerate(zip(mix, served)):
        assert s.status == 0, s.error
        r = ctx.solve_pairs(p, n)
        assert (r.kmax, r.truss_edges) == (want["kmax"][i], want["truss_edges"][i]), f"query {i} alone"
        assert (s.kmax, s.truss_edges) == (r.kmax, r.truss_edges), f"query {i} in the mix"
        edges = ctx.truss_edges()
        if len(edges):
            tagged = np.empty((len(edges), 3), dtype="<u4")
            tagged[:, 0] = i
            tagged[:, 1:] = edges
            h.update(tagged.tobytes())
    assert h.hexdigest() == want["truss_sha256"], "kmax-truss edge sets differ from the oracle's"
