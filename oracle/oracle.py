"""ORACLE — test infrastructure only (tests/, smoke(), bench.py cpu_baseline).

ctypes binding of ``kt_oracle.c`` (the CPU restatement of the reference's
k-max-truss path) plus helpers that run the reference's own ``kc_cpu`` binary
from ``oracle/_ref`` (built from the reference sources by ``oracle/Makefile``).
Nothing in the product imports this module.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
KC_CPU = HERE / "_ref" / "kc_cpu"

_u64p = ctypes.c_void_p


class KtoResult(ctypes.Structure):
    _fields_ = [("kmax", ctypes.c_uint32), ("truss_edges", ctypes.c_uint64),
                ("k_hat", ctypes.c_uint32), ("k_hat_edges", ctypes.c_uint64),
                ("max_core", ctypes.c_uint32), ("core_k", ctypes.c_uint32),
                ("n_core", ctypes.c_uint32), ("m_core", ctypes.c_uint64)]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-C", str(HERE), "liboracle.so"], check=True,
                           stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(str(LIB))
        V = ctypes.c_void_p
        L.kto_parse_tsv.restype = ctypes.c_int64
        L.kto_parse_tsv.argtypes = [V, ctypes.c_uint64, V, V, ctypes.c_uint64]
        L.kto_csr_from_lines.restype = None
        L.kto_csr_from_lines.argtypes = [ctypes.c_uint64, V, V, ctypes.c_uint32, V, V]
        L.kto_kcore.restype = ctypes.c_uint64
        L.kto_kcore.argtypes = [ctypes.c_uint32, V, V, ctypes.c_uint32, V]
        L.kto_max_core.restype = ctypes.c_uint32
        L.kto_max_core.argtypes = [ctypes.c_uint32, V, V, V]
        L.kto_rank.restype = ctypes.c_uint32
        L.kto_rank.argtypes = [ctypes.c_uint32, V, V, V, V, V, V]
        L.kto_oriented.restype = ctypes.c_uint64
        L.kto_oriented.argtypes = [ctypes.c_uint32, V, V, V, ctypes.c_uint32, V, V]
        L.kto_undirected.restype = None
        L.kto_undirected.argtypes = [ctypes.c_uint32, V, V, V, V, V, V, V]
        L.kto_support.restype = None
        L.kto_support.argtypes = [ctypes.c_uint32, V, V, V]
        L.kto_solve.restype = ctypes.c_int
        L.kto_solve.argtypes = [ctypes.c_uint32, V, V, ctypes.POINTER(KtoResult), V,
                                ctypes.c_uint64]
        L.kto_core_numbers.restype = ctypes.c_uint32
        L.kto_core_numbers.argtypes = [ctypes.c_uint32, V, V, V]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data


# ---------------------------------------------------------------- graphs ---

def lines_from_pairs(pairs: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Directed line list (x, y) in the TSV contract order: both directions,
    sorted by (y, x)."""
    pairs = np.asarray(pairs, dtype=np.uint32).reshape(-1, 2)
    x = np.concatenate([pairs[:, 0], pairs[:, 1]])
    y = np.concatenate([pairs[:, 1], pairs[:, 0]])
    order = np.lexsort((x, y))
    return np.ascontiguousarray(x[order]), np.ascontiguousarray(y[order])


def csr_from_lines(xs: np.ndarray, ys: np.ndarray, n: int | None = None):
    xs = np.ascontiguousarray(xs, dtype=np.uint32)
    ys = np.ascontiguousarray(ys, dtype=np.uint32)
    if n is None:
        n = int(ys.max()) + 1 if ys.size else 0
    off = np.zeros(n + 1, dtype=np.uint64)
    adj = np.zeros(max(xs.size, 1), dtype=np.uint32)
    lib().kto_csr_from_lines(xs.size, _p(xs), _p(ys), n, _p(off), _p(adj))
    return off, adj[: xs.size], n


def csr_from_pairs(pairs: np.ndarray, n: int | None = None):
    xs, ys = lines_from_pairs(pairs)
    if n is None:
        n = int(ys.max()) + 1 if ys.size else 0
    return csr_from_lines(xs, ys, n)


def parse_tsv_bytes(data: bytes):
    L = lib()
    buf = ctypes.create_string_buffer(data, len(data))
    nl = L.kto_parse_tsv(buf, len(data), None, None, 0)
    xs = np.zeros(max(nl, 1), dtype=np.uint32)
    ys = np.zeros(max(nl, 1), dtype=np.uint32)
    L.kto_parse_tsv(buf, len(data), _p(xs), _p(ys), nl)
    return xs[:nl], ys[:nl]


# ---------------------------------------------------------------- stages ---

def kcore_removed(off, adj, n, k) -> np.ndarray:
    removed = np.zeros(max(n, 1), dtype=np.uint8)
    lib().kto_kcore(n, _p(off), _p(adj), k, _p(removed))
    return removed[:n]


def max_core(off, adj, n):
    removed = np.zeros(max(n, 1), dtype=np.uint8)
    k = lib().kto_max_core(n, _p(off), _p(adj), _p(removed))
    return int(k), removed[:n]


def core_numbers(off, adj, n) -> np.ndarray:
    core = np.zeros(max(n, 1), dtype=np.uint32)
    lib().kto_core_numbers(n, _p(off), _p(adj), _p(core))
    return core[:n]


def rank(off, adj, n, removed):
    to = np.zeros(max(n, 1), dtype=np.uint32)
    inv = np.zeros(max(n, 1), dtype=np.uint32)
    degc = np.zeros(max(n, 1), dtype=np.uint32)
    removed = np.ascontiguousarray(removed, dtype=np.uint8)
    nc = lib().kto_rank(n, _p(off), _p(adj), _p(removed), _p(to), _p(inv), _p(degc))
    return int(nc), to[:n], inv[:nc], degc[:n]


def oriented(off, adj, n, to, nc):
    off_d = np.zeros(nc + 1, dtype=np.uint64)
    m = lib().kto_oriented(n, _p(off), _p(adj), _p(to), nc, _p(off_d), None)
    adj_d = np.zeros(max(m, 1), dtype=np.uint32)
    lib().kto_oriented(n, _p(off), _p(adj), _p(to), nc, _p(off_d), _p(adj_d))
    return off_d, adj_d[:m]


def undirected(off_d, adj_d, nc):
    m = int(off_d[nc])
    off_u = np.zeros(nc + 1, dtype=np.uint64)
    adj_u = np.zeros(max(2 * m, 1), dtype=np.uint32)
    fa = np.zeros(max(2 * m, 1), dtype=np.uint64)
    lson = np.zeros(max(m, 1), dtype=np.uint64)
    rson = np.zeros(max(m, 1), dtype=np.uint64)
    adj_d = np.ascontiguousarray(adj_d, dtype=np.uint32)
    if adj_d.size == 0:
        adj_d = np.zeros(1, dtype=np.uint32)
    lib().kto_undirected(nc, _p(off_d), _p(adj_d), _p(off_u), _p(adj_u), _p(fa), _p(lson),
                         _p(rson))
    return off_u, adj_u[: 2 * m], fa[: 2 * m], lson[:m], rson[:m]


def support2(off_d, adj_d, nc) -> np.ndarray:
    m = int(off_d[nc])
    sup = np.zeros(max(m, 1), dtype=np.uint32)
    adj_d = np.ascontiguousarray(adj_d, dtype=np.uint32)
    if adj_d.size == 0:
        adj_d = np.zeros(1, dtype=np.uint32)
    lib().kto_support(nc, _p(off_d), _p(adj_d), _p(sup))
    return sup[:m]


@dataclass
class Solve:
    kmax: int
    truss_edges: int
    k_hat: int
    k_hat_edges: int
    max_core: int
    core_k: int
    n_core: int
    m_core: int
    truss: np.ndarray | None = None


def solve_csr(off, adj, n, want_edges: bool = False) -> Solve:
    r = KtoResult()
    m = int(off[n]) // 2 if n else 0
    buf = np.zeros((max(m, 1), 2), dtype=np.uint32) if want_edges else None
    adj = np.ascontiguousarray(adj, dtype=np.uint32)
    if adj.size == 0:
        adj = np.zeros(1, dtype=np.uint32)
    lib().kto_solve(n, _p(off), _p(adj), ctypes.byref(r), _p(buf) if want_edges else None,
                    max(m, 1))
    s = Solve(r.kmax, r.truss_edges, r.k_hat, r.k_hat_edges, r.max_core, r.core_k, r.n_core,
              r.m_core)
    if want_edges:
        s.truss = buf[: r.truss_edges].copy()
    return s


def solve_pairs(pairs, n=None, want_edges: bool = False) -> Solve:
    off, adj, n = csr_from_pairs(pairs, n)
    return solve_csr(off, adj, n, want_edges)


def trussness(pairs) -> dict:
    """Per-edge trussness {(u, v): t}, u < v, by the textbook peel (Wang and
    Cheng, "Truss decomposition in massive networks", PVLDB 2012): at level
    k (from 2) remove, one at a time, the edges whose support among the
    remaining edges is <= k - 2, updating their triangles' other edges; a
    removed edge has trussness k. The reference reports only kmax and |T|
    (src/kc_cpu.cpp:474-488 is its per-level delete rule): edges with
    t = kmax are its kmax-truss. Pure-Python loops: small graphs only."""
    adj: dict[int, set] = {}
    edges = set()
    for a, b in np.asarray(pairs, dtype=np.int64).reshape(-1, 2).tolist():
        if a == b:
            continue
        u, v = (a, b) if a < b else (b, a)
        edges.add((u, v))
        adj.setdefault(u, set()).add(v)
        adj.setdefault(v, set()).add(u)
    sup = {e: len(adj[e[0]] & adj[e[1]]) for e in edges}
    out = {}
    k = 2
    while sup:
        queue = [e for e, s in sup.items() if s <= k - 2]
        while queue:
            e = queue.pop()
            if e not in sup:
                continue
            u, v = e
            for w in adj[u] & adj[v]:
                for f in ((min(u, w), max(u, w)), (min(v, w), max(v, w))):
                    sup[f] -= 1
                    if sup[f] == k - 2:
                        queue.append(f)
            adj[u].discard(v)
            adj[v].discard(u)
            del sup[e]
            out[e] = k
        k += 1
    return out


# ------------------------------------------------ the reference's kc_cpu ---

_RESULT = re.compile(rb"kmax = (\d+), Edges in kmax-truss = (\d+)\.")
_EDGE_NUM = re.compile(rb"Edge Num: (\d+)")
_READ_T = re.compile(rb"Reading file Time Cost: ([0-9.]+)s")
_ALL_T = re.compile(rb"All time cost: ([0-9.]+)s")


def kc_cpu_available() -> bool:
    return KC_CPU.exists() and os.access(KC_CPU, os.X_OK)


@dataclass
class RefRun:
    kmax: int
    truss_edges: int
    edge_num: int
    t_read: float
    t_all: float
    wall: float
    threads: int

    @property
    def t_solve(self) -> float:
        return self.t_all - self.t_read


KC_GPU_LINKED = HERE / "_ref" / "kc_gpu_linked"


def run_kc_cpu(tsv: str | os.PathLike, threads: int | None = None, timeout: float = 600.0,
               binary: Path | None = None) -> RefRun:
    """Run the reference kc_cpu (or another binary built from the reference's
    host driver, e.g. kc_gpu_linked) on a TSV. Its parser sizes per-thread
    regions by omp_get_num_procs() and indexes them by omp_get_thread_num()
    (src/kc_cpu.cpp:782-805), so the process is pinned to exactly `threads`
    CPUs with OMP_NUM_THREADS equal to that count."""
    KC = binary or KC_CPU
    cpus = sorted(os.sched_getaffinity(0))
    if threads is None:
        threads = len(cpus)
    threads = max(1, min(threads, len(cpus)))
    use = set(cpus[:threads])
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    t0 = time.perf_counter()
    p = subprocess.run([str(KC), "-f", str(tsv)], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, timeout=timeout,
                       preexec_fn=lambda: os.sched_setaffinity(0, use))
    wall = time.perf_counter() - t0
    if p.returncode != 0:
        raise RuntimeError(f"kc_cpu exited {p.returncode}: {p.stderr[-400:]!r}")
    m = _RESULT.search(p.stdout)
    if not m:
        raise RuntimeError(f"kc_cpu printed no result: {p.stdout[-400:]!r}")
    en = _EDGE_NUM.search(p.stdout)
    tr = _READ_T.search(p.stdout)
    ta = _ALL_T.search(p.stdout)
    return RefRun(int(m.group(1)), int(m.group(2)), int(en.group(1)) if en else -1,
                  float(tr.group(1)) if tr else float("nan"),
                  float(ta.group(1)) if ta else float("nan"), wall, threads)
