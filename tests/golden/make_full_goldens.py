"""Writes tests/golden/full_configs.json: the oracle's answers on the
benchmark workloads at their full sizes (run in the build container; the
oracle is the CPU restatement in oracle/, pinned by tests/test_oracle.py).

  rmat22        RMAT scale 22, ef 16, (.57,.19,.19,.05), relabelled, seed 1
  chung_lu_50m  Chung-Lu gamma 2.1, n 50M, 505M samples (~500M edges after
                dedup), planted 60-clique, seed 2
  batch4096     4096 graphs, n log-uniform in [64, 65536], ER or R-MAT,
                average degree U[8, 32], clique U[4, 40], seed 3
  serve160      160 serving queries, lognormal edges (median 200k, sigma
                1.5, cap 50M), R-MAT or Chung-Lu, clique U[8, 64], seed 4

For every graph: n, m, sha256 of the generated pairs (so a GPU test can tell
a generator change from a solver change), kmax, |T| and sha256 of the
kmax-truss edge set (original ids, u < v, sorted by (u, v), uint32
little-endian). The batch and serving mixes keep per-graph (kmax, |T|) and
one digest over all their truss edge sets (graph index, u, v).

    python tests/golden/make_full_goldens.py [names...]
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from kc_cuda_amd import gen  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = HERE / "full_configs.json"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u4").tobytes()).hexdigest()


def oracle_solve(pairs: np.ndarray, n: int):
    """The oracle's two-pass solve with the kmax-truss edge set. The CSR is
    built from the line list in generation order (row y receives x); the
    solve does not depend on the order inside a row."""
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32)
    xs = np.concatenate([pairs[:, 0], pairs[:, 1]])
    ys = np.concatenate([pairs[:, 1], pairs[:, 0]])
    off, adj, n = O.csr_from_lines(xs, ys, n)
    del xs, ys
    return O.solve_csr(off, adj, n, want_edges=True)


def single(name: str, pairs: np.ndarray, n: int) -> dict:
    t0 = time.perf_counter()
    r = oracle_solve(pairs, n)
    dt = time.perf_counter() - t0
    truss = r.truss[np.lexsort((r.truss[:, 1], r.truss[:, 0]))]
    print(f"{name}: n={n} m={len(pairs)} kmax={r.kmax} |T|={r.truss_edges} "
          f"(oracle {dt:.1f}s)", flush=True)
    return dict(n=int(n), m=int(len(pairs)), pairs_sha256=sha(pairs), kmax=int(r.kmax),
                truss_edges=int(r.truss_edges), truss_sha256=sha(truss),
                k_hat=int(r.k_hat), max_core=int(r.max_core), oracle_seconds=round(dt, 1))


def mix(name: str, graphs) -> dict:
    """graphs: iterable of (pairs, n). One digest over (g, u, v) of every
    graph's sorted kmax-truss."""
    h_pairs = hashlib.sha256()
    h_truss = hashlib.sha256()
    ns, ms, kmax, cnt = [], [], [], []
    t0 = time.perf_counter()
    for g, (pairs, n) in enumerate(graphs):
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32)
        r = oracle_solve(pairs, n)
        truss = r.truss[np.lexsort((r.truss[:, 1], r.truss[:, 0]))]
        h_pairs.update(np.ascontiguousarray(pairs, dtype="<u4").tobytes())
        if len(truss):
            tagged = np.empty((len(truss), 3), dtype="<u4")
            tagged[:, 0] = g
            tagged[:, 1:] = truss
            h_truss.update(tagged.tobytes())
        ns.append(int(n))
        ms.append(int(len(pairs)))
        kmax.append(int(r.kmax))
        cnt.append(int(r.truss_edges))
        if g % 256 == 0:
            print(f"  {name}: graph {g} n={n} m={len(pairs)} kmax={r.kmax} "
                  f"({time.perf_counter() - t0:.0f}s)", flush=True)
    print(f"{name}: {len(ns)} graphs, {sum(ms)} edges ({time.perf_counter() - t0:.0f}s)",
          flush=True)
    return dict(graphs=len(ns), n=ns, m=ms, kmax=kmax, truss_edges=cnt,
                pairs_sha256=h_pairs.hexdigest(), truss_sha256=h_truss.hexdigest())


def batch_graphs():
    pairs, offs, ns = gen.batch()
    for g in range(len(ns)):
        yield pairs[int(offs[g]):int(offs[g + 1])], int(ns[g])


CASES = {
    "rmat22": lambda: single("rmat22", *gen.workload("rmat22")),
    "chung_lu_50m": lambda: single("chung_lu_50m", *gen.workload("chung_lu_50m")),
    "batch4096": lambda: mix("batch4096", batch_graphs()),
    "serve160": lambda: mix("serve160", gen.serve_mix(160, 200_000, 1.5)),
}


def main(argv: list[str]) -> None:
    names = argv or list(CASES)
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    for name in names:
        data[name] = CASES[name]()
        # one line per workload keeps the per-graph lists compact
        body = ",\n".join(f" {json.dumps(k)}: {json.dumps(data[k], sort_keys=True)}"
                          for k in sorted(data))
        OUT.write_text("{\n" + body + "\n}\n")


if __name__ == "__main__":
    main(sys.argv[1:])
