"""Writes the golden fixtures under tests/golden/ (run in the build container,
where the reference sources are available to oracle/Makefile).

  kc_cpu_results.json  (kmax, |T|) of the reference's own kc_cpu, compiled
                       from its sources into oracle/_ref, on graphs from the
                       project's generator written in the reference's TSV
                       contract. Each graph is run twice; the two runs and the
                       'Edge Num' line must agree.
  networkx_truss.json  kmax and the kmax-truss edge set from networkx's
                       k_truss on tiny random graphs (networkx is present in
                       the build container only).
  tiny/*.tsv           small TSV inputs with their expected answer lines.

    python tests/golden/make_golden.py [--large]
"""
from __future__ import annotations

import hashlib
import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from kc_cuda_amd import gen  # noqa: E402
from oracle import oracle as O  # noqa: E402
from graphs import closed_forms, make_workload, random_graph  # noqa: E402

SMALL = {
    "er16k_clique24": dict(kind="er_clique", n=16384, avg_deg=16.0, clique=24, seed=0),
    "er3k_clique17": dict(kind="er_clique", n=3000, avg_deg=10.0, clique=17, seed=11),
    "rmat12": dict(kind="rmat", scale=12, edge_factor=16, seed=1),
    "rmat14_clique12_s7": dict(kind="rmat", scale=14, edge_factor=16, clique=12, seed=7),
    "rmat16": dict(kind="rmat", scale=16, edge_factor=16, seed=1),
    "rmat18": dict(kind="rmat", scale=18, edge_factor=16, seed=1),
    "chung_lu_200k": dict(kind="chung_lu", n=200_000, gamma=2.1, samples=2_000_000, clique=30, seed=2),
}
LARGE = {
    "rmat20": dict(kind="rmat", scale=20, edge_factor=16, seed=1),
    "rmat22": dict(kind="rmat", scale=22, edge_factor=16, seed=1),
}


make = make_workload


def kc_cpu_case(name: str, params: dict, tmp: Path) -> dict:
    pairs, n = make(params)
    tsv = tmp / f"{name}.tsv"
    gen.write_tsv(tsv, pairs)
    sha = hashlib.sha256(tsv.read_bytes()).hexdigest()
    runs = []
    failures = []
    limit = 120 if len(pairs) < 5_000_000 else 900
    while len(runs) < 2:
        # kc_cpu's parser can overflow its fixed per-thread buffers when its
        # threads are scheduled unevenly (src/kc_cpu.cpp:778-805), and its
        # racy peel can livelock: retry with a bounded wall time.
        try:
            runs.append(O.run_kc_cpu(tsv, timeout=limit))
        except (RuntimeError, subprocess.TimeoutExpired) as e:
            failures.append(type(e).__name__)
            print(f"  {name}: kc_cpu run failed ({type(e).__name__}); retrying", flush=True)
            if len(failures) >= 4:
                tsv.unlink()
                ref = O.solve_pairs(pairs, n)
                print(f"{name}: kc_cpu never finished; oracle ({ref.kmax}, {ref.truss_edges})")
                return dict(params=params, n=int(n), m=int(len(pairs)), tsv_sha256=sha,
                            kmax=None, truss_edges=None, kc_cpu_failures=failures,
                            oracle_kmax=ref.kmax, oracle_truss_edges=ref.truss_edges)
    for r in runs:
        assert r.edge_num == len(pairs), (name, r.edge_num, len(pairs))
    assert (runs[0].kmax, runs[0].truss_edges) == (runs[1].kmax, runs[1].truss_edges), name
    tsv.unlink()
    print(f"{name}: m={len(pairs)} -> ({runs[0].kmax}, {runs[0].truss_edges})", flush=True)
    return dict(params=params, n=int(n), m=int(len(pairs)), tsv_sha256=sha,
                kmax=runs[0].kmax, truss_edges=runs[0].truss_edges,
                kc_cpu_t_solve_s=[round(r.t_solve, 6) for r in runs],
                kc_cpu_threads=runs[0].threads)


def networkx_cases() -> list:
    import networkx as nx
    out = []
    for seed in range(40):
        n = 8 + seed % 23
        p = 0.15 + 0.5 * ((seed * 7) % 10) / 10
        pairs = random_graph(n, p, 1000 + seed)
        g = nx.Graph()
        g.add_nodes_from(range(n))
        g.add_edges_from(map(tuple, pairs.tolist()))
        kmax, best = 0, []
        k = 3
        while True:
            t = nx.k_truss(g, k)
            if t.number_of_edges() == 0:
                break
            kmax = k
            best = sorted((min(a, b), max(a, b)) for a, b in t.edges())
            k += 1
        out.append(dict(seed=1000 + seed, n=n, p=p, pairs=pairs.tolist(), kmax=kmax,
                         truss=[list(e) for e in best]))
    return out


def tiny_tsvs(tmp: Path) -> None:
    d = HERE / "tiny"
    d.mkdir(exist_ok=True)
    expect = {}
    for name, pairs, n, _k, _c in closed_forms():
        if len(pairs) == 0:
            continue
        safe = "".join(c if c.isalnum() else "_" for c in name)
        path = d / f"{safe}.tsv"
        gen.write_tsv(path, pairs)
        r = O.run_kc_cpu(path, timeout=60)
        expect[path.name] = f"kmax = {r.kmax}, Edges in kmax-truss = {r.truss_edges}."
    (d / "expected.json").write_text(json.dumps(expect, indent=1, sort_keys=True) + "\n")


def main() -> None:
    if not O.kc_cpu_available():
        raise SystemExit("oracle/_ref/kc_cpu is not built (make -C oracle ref)")
    large = "--large" in sys.argv
    path = HERE / "kc_cpu_results.json"
    res = json.loads(path.read_text()) if path.exists() else {}
    with tempfile.TemporaryDirectory() as td:
        tmp = Path(td)
        cases = dict(SMALL)
        if large:
            cases.update(LARGE)
        for name, params in cases.items():
            if name in res and res[name]["params"] == params:
                continue
            res[name] = kc_cpu_case(name, params, tmp)
            path.write_text(json.dumps(res, indent=1, sort_keys=True) + "\n")
        tiny_tsvs(tmp)
    (HERE / "networkx_truss.json").write_text(json.dumps(networkx_cases()) + "\n")


if __name__ == "__main__":
    main()
