"""Synthetic graph workloads for the k-max-truss path.

Thin ctypes wrapper over ``csrc/ktgen.cpp`` (built into ``lib/libktgen.so``).
Every generator returns an ``(m, 2)`` uint32 array of canonical undirected
pairs ``u < v`` sorted by ``(u, v)`` (no self-loops, no duplicates) and the
vertex count ``n``. ``write_tsv`` emits the input contract of the
reference's ``kc_gpu -f`` (src/kc_gpu.cpp:243-312): both directions, grouped
by the second column ascending.

The named workloads mirror the benchmark configurations:

* ``er_clique``   ER G(16384, avg deg 16) + 24-clique, seed 0
* ``rmat``        RMAT scale 22, ef 16, (.57,.19,.19,.05), relabelled, seed 1
* ``chung_lu``    power-law gamma 2.1, n 50M, ~500M edges + 60-clique, seed 2
* ``batch``       4096 graphs, n log-uniform in [64, 65536], ER or R-MAT,
                  average degree U[8, 32], clique U[4, 40], seed 3
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib

_u32pp = ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32))


def _gen():
    lib = _lib.load_host("libktgen.so")
    if not getattr(lib, "_kt_typed", False):
        lib.ktgen_er_clique.restype = ctypes.c_int64
        lib.ktgen_er_clique.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32,
                                        ctypes.c_uint64, _u32pp]
        lib.ktgen_rmat.restype = ctypes.c_int64
        lib.ktgen_rmat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                                   _u32pp]
        lib.ktgen_chung_lu.restype = ctypes.c_int64
        lib.ktgen_chung_lu.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_int64, ctypes.c_int,
                                       ctypes.c_uint32, ctypes.c_uint64, _u32pp]
        lib.ktgen_batch.restype = ctypes.c_int64
        lib.ktgen_batch.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                    _u32pp, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint64)), _u32pp]
        lib.ktgen_write_tsv.restype = ctypes.c_int64
        lib.ktgen_write_tsv.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64]
        lib.ktgen_free.argtypes = [ctypes.c_void_p]
        lib.ktgen_free.restype = None
        lib._kt_typed = True
    return lib


def _take(lib, m: int, ptr) -> np.ndarray:
    if m < 0:
        raise MemoryError("graph generator ran out of memory")
    try:
        if m == 0:
            return np.zeros((0, 2), dtype=np.uint32)
        view = np.ctypeslib.as_array(ptr, shape=(m * 2,))
        return view.copy().reshape(m, 2)
    finally:
        lib.ktgen_free(ptr)


def er_clique(n: int = 16384, avg_deg: float = 16.0, clique: int = 24, seed: int = 0):
    lib = _gen()
    p = ctypes.POINTER(ctypes.c_uint32)()
    m = lib.ktgen_er_clique(n, avg_deg, clique, seed, ctypes.byref(p))
    return _take(lib, m, p), n


def rmat(scale: int = 22, edge_factor: int = 16, a: float = 0.57, b: float = 0.19,
         c: float = 0.19, relabel: bool = True, clique: int = 0, seed: int = 1):
    lib = _gen()
    p = ctypes.POINTER(ctypes.c_uint32)()
    m = lib.ktgen_rmat(scale, edge_factor, a, b, c, int(relabel), clique, seed, ctypes.byref(p))
    return _take(lib, m, p), 1 << scale


def chung_lu(n: int = 50_000_000, gamma: float = 2.1, samples: int = 505_000_000,
             clique: int = 60, seed: int = 2, relabel: bool = True):
    lib = _gen()
    p = ctypes.POINTER(ctypes.c_uint32)()
    m = lib.ktgen_chung_lu(n, gamma, samples, int(relabel), clique, seed, ctypes.byref(p))
    return _take(lib, m, p), n


def batch(count: int = 4096, n_min: int = 64, n_max: int = 65536, deg_min: float = 8.0,
          deg_max: float = 32.0, clique_min: int = 4, clique_max: int = 40, seed: int = 3):
    """A batch of independent graphs (ER or R-MAT, planted cliques).

    Returns (pairs, offsets, ns): pairs (M, 2) uint32 canonical and sorted
    within each graph, graphs concatenated; offsets (count + 1) uint64 pair
    offsets; ns (count,) uint32 vertex counts."""
    lib = _gen()
    p = ctypes.POINTER(ctypes.c_uint32)()
    o = ctypes.POINTER(ctypes.c_uint64)()
    q = ctypes.POINTER(ctypes.c_uint32)()
    m = lib.ktgen_batch(count, n_min, n_max, deg_min, deg_max, clique_min, clique_max, seed,
                        ctypes.byref(p), ctypes.byref(o), ctypes.byref(q))
    if m < 0:
        raise MemoryError("graph generator ran out of memory")
    try:
        offs = np.ctypeslib.as_array(o, shape=(count + 1,)).copy()
        ns = np.ctypeslib.as_array(q, shape=(count,)).copy() if count else np.zeros(0, np.uint32)
    finally:
        lib.ktgen_free(o)
        lib.ktgen_free(q)
    return _take(lib, m, p), offs, ns


def serve_query(target_edges: int, kind: str, clique: int, seed: int):
    """One serving query of about target_edges edges: R-MAT (edge factor 16,
    (.57,.19,.19,.05), scale chosen so that 15.3 * 2^scale is nearest the
    target: that many edges survive deduplication) or Chung-Lu (gamma 2.1,
    average degree 10, as many samples as target edges), with a planted
    clique. Returns (pairs, n)."""
    if kind == "rmat":
        scale = int(min(28, max(5, round(np.log2(max(target_edges, 1) / 15.3)))))
        return rmat(scale, 16, 0.57, 0.19, 0.19, True, clique, seed)
    n = max(2 * clique, target_edges // 5)
    return chung_lu(n, 2.1, max(target_edges, clique * clique), clique, seed)


def serve_mix(count: int = 256, median: float = 200_000, sigma: float = 1.5,
              cap: int = 50_000_000, seed: int = 4):
    """Serving queries: edge counts lognormal (median, sigma) capped at cap,
    each R-MAT or Chung-Lu (50/50) with a planted clique of U[8, 64]
    vertices. Returns a list of (pairs, n) in arrival order."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        target = int(min(cap, max(1000, rng.lognormal(np.log(median), sigma))))
        kind = "rmat" if rng.random() < 0.5 else "chung_lu"
        clique = int(rng.integers(8, 65))
        out.append(serve_query(target, kind, clique, int(rng.integers(1, 2**31 - 1))))
    return out


def write_tsv(path: str | os.PathLike, pairs: np.ndarray) -> int:
    """Write pairs in the reference's TSV contract; returns bytes written."""
    lib = _gen()
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32)
    r = lib.ktgen_write_tsv(os.fsencode(str(path)), pairs.ctypes.data, pairs.shape[0])
    if r < 0:
        raise OSError(f"cannot write {path}")
    return int(r)


def workload(name: str):
    """Named benchmark workloads -> (pairs, n)."""
    if name == "er16k_clique24":
        return er_clique(16384, 16.0, 24, 0)
    if name == "rmat22":
        return rmat(22, 16, 0.57, 0.19, 0.19, True, 0, 1)
    if name == "chung_lu_50m":
        return chung_lu(50_000_000, 2.1, 505_000_000, 60, 2)
    if name.startswith("rmat"):
        scale = int(name[4:])
        return rmat(scale, 16, 0.57, 0.19, 0.19, True, 0, 1)
    raise KeyError(name)
