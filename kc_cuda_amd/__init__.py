"""kc_cuda_amd — MI355X-native k-max-truss (drop-in for caixiaofan/KC_CUDA's kc_gpu path).

The compute path is ``lib/libkc_amd.so`` (HIP for gfx950 + C++ host code),
reached through its C ABI (``include/ktruss.h``). This module is a thin ctypes
binding over it; it has no CPU fallback and raises when the library is absent.

    from kc_cuda_amd import Context
    with Context() as ctx:
        res = ctx.solve_pairs(pairs, n)       # pairs: (m, 2) uint32, u != v
        print(res.kmax, res.truss_edges)

``kmax_truss(pairs, n)`` mirrors the reference's ``kmax_truss`` two-pass
driver (src/kc_gpu.cpp:129-208, 336-341) and returns ``(kmax, edges)``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import NativeLibraryMissing  # noqa: F401

__all__ = ["Context", "KtResult", "KtError", "Server", "ServeResult", "kmax_truss", "native_library",
           "parse_tsv", "NativeLibraryMissing", "EXPORTS",
           "solve_device_noalloc"]

# Every symbol declared in include/ktruss.h and include/kc_gpu_c.h.
EXPORTS = [
    "kt_last_error", "kt_device_count", "kt_create", "kt_destroy", "kt_malloc", "kt_free",
    "kt_memcpy_h2d", "kt_memcpy_d2h", "kt_memcpy_d2d", "kt_synchronize", "kt_solve_pairs_device",
    "kt_solve_pairs_host", "kt_solve_lines_host", "kt_solve_tsv", "kt_truss_edges", "kt_trussness",
    "kt_push_probe", "kt_push_stats_get", "kt_kernel_probe",
    "kt_parse_tsv", "kt_parse_tsv_device", "kt_set_tuning", "kt_last_csr", "kt_last_stats", "kt_batch_solve",
    "kt_batch_autotune", "kt_serve_create", "kt_serve_destroy", "kt_serve_submit", "kt_serve_wait",
    "kt_serve_counters_get", "kt_serve_footprint", "kt_solve_pairs_device_noalloc",
]


class KtError(RuntimeError):
    pass


class KtResult(ctypes.Structure):
    _fields_ = [
        ("kmax", ctypes.c_uint32), ("max_core", ctypes.c_uint32),
        ("truss_edges", ctypes.c_uint64), ("n", ctypes.c_uint32), ("k_hat", ctypes.c_uint32),
        ("m", ctypes.c_uint64), ("k_hat_edges", ctypes.c_uint64), ("core_k", ctypes.c_uint32),
        ("n_core", ctypes.c_uint32), ("m_core", ctypes.c_uint64),
        ("peel_rounds", ctypes.c_uint64), ("levels", ctypes.c_uint32), ("lanes", ctypes.c_uint32),
        ("ms_total", ctypes.c_float), ("ms_build", ctypes.c_float), ("ms_kcore", ctypes.c_float),
        ("ms_pass1", ctypes.c_float), ("ms_extract", ctypes.c_float), ("ms_count", ctypes.c_float),
        ("ms_peel", ctypes.c_float), ("ms_count_kernel", ctypes.c_float),
        ("compactions", ctypes.c_uint32), ("recounts", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class KtStats(ctypes.Structure):
    _fields_ = [("n_core", ctypes.c_uint32), ("max_out_degree", ctypes.c_uint32),
                ("m_core", ctypes.c_uint64), ("w_core", ctypes.c_uint64),
                ("w_peeled", ctypes.c_uint64), ("truss_edges", ctypes.c_uint64),
                ("count_kernel", ctypes.c_uint32), ("core_words", ctypes.c_uint32),
                ("vertex_rounds", ctypes.c_uint32), ("pad", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class KtServeResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("slot", ctypes.c_uint32), ("result", KtResult),
                ("footprint", ctypes.c_uint64), ("charged", ctypes.c_uint64),
                ("slot_bytes", ctypes.c_uint64), ("queued_ms", ctypes.c_double), ("run_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("group", ctypes.c_uint32)]


class KtBatchTuning(ctypes.Structure):
    _fields_ = [("count_lanes", ctypes.c_int32), ("peel_lanes", ctypes.c_int32),
                ("edge_count", ctypes.c_int32), ("trials", ctypes.c_uint32),
                ("drop_ratio", ctypes.c_double), ("ms_default", ctypes.c_double),
                ("ms_best", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class KtPushStats(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64), ("slots", ctypes.c_uint64), ("vertices", ctypes.c_uint64),
                ("ms", ctypes.c_double)]


class KtServeCounters(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "submitted", "completed", "rejected", "failed", "in_flight", "peak_in_flight",
        "bytes_in_use", "peak_bytes_in_use", "over_estimate", "trims", "grouped")]

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


_V = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64


def native_library() -> ctypes.CDLL:
    lib = _lib.load_device()
    if getattr(lib, "_kt_typed", False):
        return lib
    lib.kt_last_error.restype = ctypes.c_char_p
    lib.kt_last_error.argtypes = []
    sig = {
        "kt_device_count": [ctypes.POINTER(ctypes.c_int)],
        "kt_create": [ctypes.c_int, ctypes.POINTER(_V)],
        "kt_destroy": [_V],
        "kt_malloc": [_V, _U64, ctypes.POINTER(_V)],
        "kt_free": [_V, _V],
        "kt_memcpy_h2d": [_V, _V, _V, _U64],
        "kt_memcpy_d2h": [_V, _V, _V, _U64],
        "kt_memcpy_d2d": [_V, _V, _V, _U64],
        "kt_synchronize": [_V],
        "kt_solve_pairs_device": [_V, _V, _U64, _U32, ctypes.POINTER(KtResult)],
        "kt_solve_pairs_host": [_V, _V, _U64, _U32, ctypes.POINTER(KtResult)],
        "kt_solve_lines_host": [_V, _V, _V, _U64, _U32, ctypes.POINTER(KtResult)],
        "kt_solve_tsv": [_V, ctypes.c_char_p, ctypes.POINTER(KtResult),
                         ctypes.POINTER(ctypes.c_double)],
        "kt_truss_edges": [_V, _V, _U64, ctypes.POINTER(_U64)],
        "kt_parse_tsv": [ctypes.c_char_p, _V, _V, _U64, ctypes.POINTER(_U64),
                         ctypes.POINTER(_U32)],
        "kt_set_tuning": [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "kt_last_csr": [_V, _V, _V],
        "kt_last_stats": [_V, ctypes.POINTER(KtStats)],
        "kt_parse_tsv_device": [_V, ctypes.c_char_p, _V, _V, _U64, ctypes.POINTER(_U64), ctypes.POINTER(_U32),
                                ctypes.POINTER(ctypes.c_int)],
        "kt_trussness": [_V, _V, _U64, ctypes.POINTER(_U64)],
        "kt_push_probe": [_V, ctypes.c_int],
        "kt_push_stats_get": [_V, ctypes.POINTER(KtPushStats)],
        "kt_kernel_probe": [_V, ctypes.c_int, ctypes.c_int],
        "kt_batch_solve": [_V, _V, _V, _V, _U32, _V, _V, _V, _V, ctypes.c_int],
        "kt_batch_autotune": [_V, _V, _V, _V, _U32, _U32, _V, _V, _V, ctypes.POINTER(KtBatchTuning)],
        "kt_serve_create": [ctypes.c_int, ctypes.c_int, _U64, ctypes.POINTER(_V)],
        "kt_serve_destroy": [_V],
        "kt_serve_submit": [_V, _V, _U64, _U32, ctypes.c_int, ctypes.POINTER(_U64)],
        "kt_serve_wait": [_V, _U64, ctypes.POINTER(KtServeResult)],
        "kt_serve_counters_get": [_V, ctypes.POINTER(KtServeCounters)],
        "kt_solve_pairs_device_noalloc": [ctypes.c_int, _V, _U64, _U32, _V, _U64, ctypes.POINTER(_U64),
                                          ctypes.POINTER(KtResult)],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.kt_serve_footprint.argtypes = [_U64, _U32]
    lib.kt_serve_footprint.restype = _U64
    lib._kt_typed = True
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = native_library().kt_last_error()
        raise KtError(f"status {rc}: {msg.decode() if msg else ''}")


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def parse_tsv(path: str | os.PathLike) -> tuple[np.ndarray, np.ndarray, int]:
    """Parse an edge-list file with the reference's input contract;
    returns (xs, ys, n) where line i puts xs[i] into row ys[i]."""
    lib = native_library()
    lines = _U64()
    n = _U32()
    _check(lib.kt_parse_tsv(os.fsencode(str(path)), None, None, 0, ctypes.byref(lines),
                            ctypes.byref(n)))
    L = lines.value
    xs = np.zeros(max(L, 1), dtype=np.uint32)
    ys = np.zeros(max(L, 1), dtype=np.uint32)
    _check(lib.kt_parse_tsv(os.fsencode(str(path)), xs.ctypes.data, ys.ctypes.data, L,
                            ctypes.byref(lines), ctypes.byref(n)))
    return xs[:L], ys[:L], n.value


def solve_device_noalloc(device: int, dpairs_ptr: int, m: int, n: int, dtruss_ptr: int = 0,
                         cap: int = 0) -> tuple[KtResult, int]:
    """The allocation-free solve (kt_solve_pairs_device_noalloc): m canonical
    pairs at device address dpairs_ptr; when dtruss_ptr is non-zero, at most
    cap kmax-truss pairs are written there (device memory). Returns the
    result and the truss size."""
    lib = native_library()
    r = KtResult()
    cnt = _U64()
    _check(lib.kt_solve_pairs_device_noalloc(int(device), dpairs_ptr or None, int(m), int(n),
                                             dtruss_ptr or None, int(cap), ctypes.byref(cnt), ctypes.byref(r)))
    return r, int(cnt.value)


@dataclass
class DeviceArray:
    ctx: "Context"
    ptr: int
    nbytes: int

    def free(self) -> None:
        if self.ptr:
            _check(native_library().kt_free(self.ctx._h, self.ptr))
            self.ptr = 0


class Context:
    """One solver context: a HIP stream plus device buffers reused across
    solves (graphs stay resident until the next solve on this context)."""

    def __init__(self, device: int = 0):
        self._lib = native_library()
        h = _V()
        _check(self._lib.kt_create(device, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(self._lib.kt_destroy(self._h))
            self._h = None

    def __enter__(self) -> "Context":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    # -- device memory ------------------------------------------------------
    def to_device(self, a: np.ndarray) -> DeviceArray:
        a = np.ascontiguousarray(a)
        p = _V()
        _check(self._lib.kt_malloc(self._h, max(a.nbytes, 1), ctypes.byref(p)))
        d = DeviceArray(self, p.value, a.nbytes)
        if a.nbytes:
            _check(self._lib.kt_memcpy_h2d(self._h, p.value, a.ctypes.data, a.nbytes))
        return d

    # -- solves -------------------------------------------------------------
    def tune(self, count_lanes: int = 0, peel_lanes: int = 0, force_edge_count: bool = False,
             force_atomic_build: bool = False, core: str = "auto", pass1: str = "auto",
             vertex: str = "auto", unsorted_build: bool = False):
        """core: "auto", "sparse" (sorted rows) or "dense" (bit-matrix rows);
        pass1 on bit-matrix cores: "auto" (walk 16 levels, then probe
        thresholds), "walk" (every level) or "probe" (probe after level 1);
        vertex: the vertex peel as one "cascade" launch per threshold, as
        synchronous "rounds", or "auto" (the cascade on levels of up to
        2^24 row slots); unsorted_build: the bucketed CSR build even for
        inputs sorted by row."""
        flags = (1 if force_edge_count else 0) | (2 if force_atomic_build else 0)
        flags |= {"auto": 0, "sparse": 4, "dense": 8}[core]
        flags |= {"auto": 0, "walk": 16, "probe": 32}[pass1]
        flags |= {"auto": 0, "rounds": 64, "cascade": 128}[vertex]
        flags |= 256 if unsorted_build else 0
        _check(self._lib.kt_set_tuning(self._h, count_lanes, peel_lanes, flags))

    def solve_pairs(self, pairs: np.ndarray, n: int) -> KtResult:
        pairs = _u32(pairs).reshape(-1, 2)
        r = KtResult()
        _check(self._lib.kt_solve_pairs_host(self._h, pairs.ctypes.data if pairs.size else None,
                                             pairs.shape[0], n, ctypes.byref(r)))
        return r

    def solve_device(self, dpairs: DeviceArray, m: int, n: int) -> KtResult:
        r = KtResult()
        _check(self._lib.kt_solve_pairs_device(self._h, dpairs.ptr, m, n, ctypes.byref(r)))
        return r

    def solve_lines(self, xs: np.ndarray, ys: np.ndarray, n: int) -> KtResult:
        xs, ys = _u32(xs), _u32(ys)
        r = KtResult()
        _check(self._lib.kt_solve_lines_host(self._h, xs.ctypes.data if xs.size else None,
                                             ys.ctypes.data if ys.size else None, xs.size, n,
                                             ctypes.byref(r)))
        return r

    def parse_tsv_device(self, path: str | os.PathLike) -> tuple[np.ndarray, np.ndarray, int, bool]:
        """The device-side parse of an edge-list file: (xs, ys, n, on_device);
        on_device is False when the host parser took the file."""
        lines, n, dev = _U64(), _U32(), ctypes.c_int()
        _check(self._lib.kt_parse_tsv_device(self._h, os.fsencode(str(path)), None, None, 0,
                                             ctypes.byref(lines), ctypes.byref(n), ctypes.byref(dev)))
        L = lines.value
        xs = np.zeros(max(L, 1), dtype=np.uint32)
        ys = np.zeros(max(L, 1), dtype=np.uint32)
        _check(self._lib.kt_parse_tsv_device(self._h, os.fsencode(str(path)), xs.ctypes.data, ys.ctypes.data,
                                             L, ctypes.byref(lines), ctypes.byref(n), ctypes.byref(dev)))
        return xs[:L], ys[:L], n.value, bool(dev.value)

    def solve_tsv(self, path: str | os.PathLike) -> tuple[KtResult, float]:
        r = KtResult()
        ms = ctypes.c_double()
        _check(self._lib.kt_solve_tsv(self._h, os.fsencode(str(path)), ctypes.byref(r),
                                      ctypes.byref(ms)))
        return r, ms.value

    def truss_edges(self) -> np.ndarray:
        """Edges of the kmax-truss of the last solve, original ids, u < v,
        sorted by (u, v)."""
        cnt = _U64()
        _check(self._lib.kt_truss_edges(self._h, None, 0, ctypes.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), 2), dtype=np.uint32)
        _check(self._lib.kt_truss_edges(self._h, out.ctypes.data, cnt.value, ctypes.byref(cnt)))
        out = out[: cnt.value]
        order = np.lexsort((out[:, 1], out[:, 0]))
        return out[order]

    def push_probe(self, mode: int) -> None:
        """Measurement of the vertex-peel push kernel: 1 stamps events at
        each launch's start and end, 2 counts its slots and vertices, 0 off."""
        _check(self._lib.kt_push_probe(self._h, int(mode)))

    def kernel_probe(self, kernel: str, mode: int) -> None:
        """Measurement of "vertex_peel" (as push_probe) or "dense_peel" (the
        bit-matrix edge-peel round; push_stats then gives slots = the
        frontier edges' d+(u) + d+(v) summed, vertices = frontier edges)."""
        _check(self._lib.kt_kernel_probe(self._h, {"vertex_peel": 0, "dense_peel": 1}[kernel], int(mode)))

    def push_stats(self) -> dict:
        s = KtPushStats()
        _check(self._lib.kt_push_stats_get(self._h, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in s._fields_}

    def trussness(self) -> np.ndarray:
        """(u, v, t) rows, u < v, sorted, for every edge of the last solved
        graph with trussness t >= 3 (the others have 2). The decomposition
        runs once (the second C call only copies); truss_edges() still
        returns the last solve's kmax-truss afterwards."""
        cnt = _U64()
        _check(self._lib.kt_trussness(self._h, None, 0, ctypes.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), 3), dtype=np.uint32)
        _check(self._lib.kt_trussness(self._h, out.ctypes.data, cnt.value, ctypes.byref(cnt)))
        out = out[: cnt.value]
        return out[np.lexsort((out[:, 1], out[:, 0]))]

    def stats(self) -> KtStats:
        """Byte-model terms W(C+), W(C+ minus T+) of the last solve."""
        s = KtStats()
        _check(self._lib.kt_last_stats(self._h, ctypes.byref(s)))
        return s

    def last_csr(self, n: int, m2: int) -> tuple[np.ndarray, np.ndarray]:
        off = np.zeros(n + 1, dtype=np.uint64)
        adj = np.zeros(max(m2, 1), dtype=np.uint32)
        _check(self._lib.kt_last_csr(self._h, off.ctypes.data, adj.ctypes.data))
        return off, adj[:m2]

    def synchronize(self) -> None:
        _check(self._lib.kt_synchronize(self._h))

    def copy_device(self, dst: DeviceArray, src: DeviceArray, nbytes: int) -> None:
        _check(self._lib.kt_memcpy_d2d(self._h, dst.ptr, src.ptr, nbytes))

    def batch_solve(self, dpairs: DeviceArray, offsets: np.ndarray, ns: np.ndarray,
                    keep: bool = False) -> "BatchResult":
        """Solve every graph of a batch (pairs already in HBM, canonical and
        sorted within each graph; offsets: count + 1 pair offsets; ns: vertex
        counts): the valid graphs are solved together as one disjoint union."""
        run = self.batch(dpairs, offsets, ns, keep)
        try:
            run.solve()
            return run.result()
        finally:
            run.free()

    def batch(self, dpairs: DeviceArray, offsets: np.ndarray, ns: np.ndarray,
              keep: bool = False) -> "BatchRun":
        """A reusable batch: result buffers stay allocated across solves."""
        return BatchRun(self, dpairs, offsets, ns, keep)

    def batch_autotune(self, dpairs: DeviceArray, offsets: np.ndarray, ns: np.ndarray,
                       sample_every: int = 1) -> tuple["BatchResult", dict]:
        """Times the batched solve of every sample_every-th graph under
        candidate settings (peel lanes, count lanes, count kernel, drop
        ratio), keeps the fastest on this context and returns the whole
        batch's results with the chosen settings and timings."""
        out = _BatchBuffers(self, offsets, ns, False)
        t = KtBatchTuning()
        try:
            _check(self._lib.kt_batch_autotune(self._h, dpairs.ptr, out.offsets.ctypes.data,
                                               out.ns.ctypes.data, out.count, sample_every,
                                               out.kmax.ptr, out.edges.ptr, out.status.ptr,
                                               ctypes.byref(t)))
            return out.fetch(), t.as_dict()
        finally:
            out.free()

    def alloc(self, nbytes: int) -> DeviceArray:
        p = _V()
        _check(self._lib.kt_malloc(self._h, max(nbytes, 1), ctypes.byref(p)))
        return DeviceArray(self, p.value, nbytes)


@dataclass
class BatchResult:
    kmax: np.ndarray    # per graph
    edges: np.ndarray   # per graph: edges in its kmax-truss
    status: np.ndarray  # per graph: 0 ok, 1 invalid input
    keep: np.ndarray | None  # per pair: 1 when in its graph's kmax-truss


class _BatchBuffers:
    def __init__(self, ctx: "Context", offsets, ns, keep: bool):
        self.ctx = ctx
        self.dpairs = None
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.ns = np.ascontiguousarray(ns, dtype=np.uint32)
        self.count = int(self.ns.shape[0])
        if self.offsets.shape[0] != self.count + 1:
            raise ValueError("offsets must have count + 1 entries")
        self.kmax = ctx.alloc(4 * self.count)
        self.edges = ctx.alloc(8 * self.count)
        self.status = ctx.alloc(4 * self.count)
        self.keep = ctx.alloc(int(self.offsets[-1])) if keep else None

    def fetch(self) -> BatchResult:
        lib = self.ctx._lib
        k = np.zeros(self.count, dtype=np.uint32)
        e = np.zeros(self.count, dtype=np.uint64)
        s = np.zeros(self.count, dtype=np.uint32)
        for arr, d in ((k, self.kmax), (e, self.edges), (s, self.status)):
            if arr.nbytes:
                _check(lib.kt_memcpy_d2h(self.ctx._h, arr.ctypes.data, d.ptr, arr.nbytes))
        kp = None
        if self.keep is not None:
            kp = np.zeros(int(self.offsets[-1]), dtype=np.uint8)
            if kp.nbytes:
                _check(lib.kt_memcpy_d2h(self.ctx._h, kp.ctypes.data, self.keep.ptr, kp.nbytes))
        return BatchResult(k, e, s, kp)

    def free(self) -> None:
        for d in (self.kmax, self.edges, self.status, self.keep):
            if d is not None:
                d.free()


class BatchRun(_BatchBuffers):
    """Device result buffers of one batch, reused across solves."""

    def __init__(self, ctx: "Context", dpairs: DeviceArray, offsets, ns, keep: bool):
        super().__init__(ctx, offsets, ns, keep)
        self.dpairs = dpairs

    def solve(self) -> None:
        lib = self.ctx._lib
        _check(lib.kt_batch_solve(self.ctx._h, self.dpairs.ptr, self.offsets.ctypes.data,
                                  self.ns.ctypes.data, self.count, self.kmax.ptr, self.edges.ptr,
                                  self.status.ptr, self.keep.ptr if self.keep else None, 0))

    def result(self) -> BatchResult:
        return self.fetch()


@dataclass
class ServeResult:
    ticket: int
    status: int          # 0 solved, 1 rejected by admission, 2 solve failed
    slot: int
    result: KtResult
    footprint: int
    charged: int
    slot_bytes: int
    queued_ms: float
    run_ms: float
    total_ms: float
    error: str
    group: int = 1       # queries solved with it as one disjoint union (1: alone)

    @property
    def kmax(self) -> int:
        return int(self.result.kmax)

    @property
    def truss_edges(self) -> int:
        return int(self.result.truss_edges)


class Server:
    """Serving loop over the native library: graphs are queued first come,
    first served and solved concurrently on `slots` solver contexts (one HIP
    stream and host thread each); a query is admitted once a slot is free and
    its charge (what the slot holds plus the query's footprint) fits in
    `hbm_budget` bytes beside the other slots' charges. When more small
    queries wait than idle slots can take, a slot takes several and solves
    them as one disjoint union (continuous batching; ServeResult.group).

        with Server(slots=8, hbm_budget=64 << 30) as srv:
            t = srv.submit(pairs, n)
            r = srv.wait(t)          # ServeResult: kmax, truss_edges, latencies
    """

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget: int = 64 << 30):
        self._lib = native_library()
        h = _V()
        _check(self._lib.kt_serve_create(device, slots, hbm_budget, ctypes.byref(h)))
        self._h = h
        self.slots = slots
        self.hbm_budget = hbm_budget
        self._held = {}  # ticket -> input kept alive until waited for

    @staticmethod
    def footprint(m: int, n: int) -> int:
        return int(native_library().kt_serve_footprint(m, n))

    def submit(self, pairs, n: int) -> int:
        """pairs: (m, 2) uint32 canonical pairs on the host, or a DeviceArray of
        them (then pass m via pairs.nbytes // 8)."""
        t = _U64()
        if isinstance(pairs, DeviceArray):
            m = pairs.nbytes // 8
            _check(self._lib.kt_serve_submit(self._h, pairs.ptr, m, n, 1, ctypes.byref(t)))
        else:
            p = _u32(pairs).reshape(-1, 2)
            m = p.shape[0]
            _check(self._lib.kt_serve_submit(self._h, p.ctypes.data if m else None, m, n, 0,
                                             ctypes.byref(t)))
        self._held[t.value] = pairs if isinstance(pairs, DeviceArray) else p
        return t.value

    def wait(self, ticket: int) -> ServeResult:
        out = KtServeResult()
        rc = self._lib.kt_serve_wait(self._h, ticket, ctypes.byref(out))
        self._held.pop(ticket, None)
        _check(rc)
        msg = self._lib.kt_last_error() if out.status != 0 else None
        return ServeResult(ticket, int(out.status), int(out.slot), out.result, int(out.footprint),
                           int(out.charged), int(out.slot_bytes), float(out.queued_ms), float(out.run_ms),
                           float(out.total_ms), msg.decode() if msg else "", int(out.group))

    def counters(self) -> dict:
        c = KtServeCounters()
        _check(self._lib.kt_serve_counters_get(self._h, ctypes.byref(c)))
        return c.as_dict()

    def close(self) -> None:
        if self._h:
            _check(self._lib.kt_serve_destroy(self._h))
            self._h = None
            self._held.clear()

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()


def kmax_truss(pairs: np.ndarray, n: int, device: int = 0) -> tuple[int, int]:
    """(kmax, edges in the kmax-truss) of an undirected simple graph."""
    with Context(device) as ctx:
        r = ctx.solve_pairs(pairs, n)
        return int(r.kmax), int(r.truss_edges)
