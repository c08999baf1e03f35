"""Python mirror of the reference's host driver over the kc_gpu.h stage
functions (C-ABI twins in include/kc_gpu_c.h).

The functions keep the reference's names and argument meanings; arrays are
numpy uint32. ``kmax_truss`` follows src/kc_gpu.cpp:68-208 (construct_Graph
then the search) and ``solve`` its two passes (src/kc_gpu.cpp:336-341).
Errors raise KtError (the C++ ABI exits instead, like the reference).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import KtError, _lib

K_MIDDLE = 19
NONE = 0xFFFFFFFF


class Link(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("t", ctypes.c_uint32), ("d", ctypes.c_uint32)]


_V = ctypes.c_void_p
_U = ctypes.c_uint32


def _lib_c():
    lib = _lib.load_device()
    if not getattr(lib, "_kc_typed", False):
        P = ctypes.POINTER(_U)
        lib.kc_last_error.restype = ctypes.c_char_p
        sig = {
            "kc_initGPU": [_U, _U],
            "kc_kmax_core_gpu": [_U, _V, _V, _V, _V, _U, P],
            "kc_count_deg_with_mask": [_V, _V, _U, _V, _V, _V, _V, P, _V, _V],
            "kc_build_nbr_direct": [_V, _V, _U, _V, _V, _V, _V, _U],
            "kc_build_graph_from_direct": [_V, _V, _U, _V, _V, _V, _V],
            "kc_tricount_direct": [_V, _V, _U, _V, _V, _V, _V, _U],
            "kc_kmax_truss_GPU": [P, P, _U, _U, _U, _U, _V, _V],
        }
        for name, args in sig.items():
            getattr(lib, name).argtypes = args
            getattr(lib, name).restype = ctypes.c_int
        lib.kc_test_cuda.restype = None
        lib._kc_typed = True
    return lib


def _ok(rc: int) -> None:
    if rc:
        raise KtError(f"status {rc}: {_lib_c().kc_last_error().decode()}")


def _p(a):
    return None if a is None else a.ctypes.data


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def initGPU(edge_num: int, N: int) -> None:
    _ok(_lib_c().kc_initGPU(edge_num, N))


def kmax_core_gpu(N, nbr_u, nbr, deg_origin, mask, launch_start):
    """mask (in/out, zero-initialised by the caller) gets 1 outside the core;
    returns k - 1 of the max-core (launch_start = 0xFFFFFFFF) or
    launch_start - 1."""
    r = _U()
    _ok(_lib_c().kc_kmax_core_gpu(N, _p(_u32(nbr_u)), _p(_u32(nbr)), _p(_u32(deg_origin)), _p(mask),
                                  launch_start, ctypes.byref(r)))
    return r.value


def count_deg_with_mask(nbr, nbr_u, N, node_delete, deg_with_mask, to):
    """Returns (li, new_N, deg, deg_direct); deg_with_mask and to are
    updated in place."""
    li = (Link * max(N, 1))()
    nn = _U()
    deg = np.zeros(max(N, 1), dtype=np.uint32)
    ddir = np.zeros(max(N, 1), dtype=np.uint32)
    _ok(_lib_c().kc_count_deg_with_mask(_p(_u32(nbr)), _p(_u32(nbr_u)), N, _p(_u32(node_delete)),
                                        _p(deg_with_mask), _p(to), ctypes.addressof(li),
                                        ctypes.byref(nn), _p(deg), _p(ddir)))
    k = nn.value
    li_arr = np.array([(li[i].x, li[i].t, li[i].d) for i in range(k)], dtype=np.uint32).reshape(-1, 3)
    return li_arr, k, deg[:k], ddir[:k]


def build_nbr_direct(nbr, nbr_u, N, node_delete, to, nbr_u_direct, new_N):
    m = int(nbr_u_direct[new_N])
    out = np.zeros(max(m, 1), dtype=np.uint32)
    _ok(_lib_c().kc_build_nbr_direct(_p(_u32(nbr)), _p(_u32(nbr_u)), N, _p(_u32(node_delete)),
                                     _p(_u32(to)), _p(_u32(nbr_u_direct)), _p(out), new_N))
    return out[:m]


def build_graph_from_direct(nbr_direct, nbr_u_direct, N):
    """Returns (nbr, fa, left_son, right_son)."""
    m = int(nbr_u_direct[N])
    nbr = np.zeros(max(2 * m, 1), dtype=np.uint32)
    fa = np.zeros(max(2 * m, 1), dtype=np.uint32)
    ls = np.zeros(max(m, 1), dtype=np.uint32)
    rs = np.zeros(max(m, 1), dtype=np.uint32)
    _ok(_lib_c().kc_build_graph_from_direct(_p(_u32(nbr_direct)), _p(_u32(nbr_u_direct)), N, _p(nbr),
                                            _p(fa), _p(ls), _p(rs)))
    return nbr[: 2 * m], fa[: 2 * m], ls[:m], rs[:m]


def tricount_direct(nbr_u_direct, nbr_direct, N, left_son=None, right_son=None, tc_limit=0):
    """Returns (nbr_tc, nbr_tc_double): doubled supports per oriented edge,
    and per undirected slot when tc_limit == 0 and the son arrays are given."""
    m = int(nbr_u_direct[N])
    tc = np.zeros(max(m, 1), dtype=np.uint32)
    tcd = np.zeros(max(2 * m, 1), dtype=np.uint32) if left_son is not None else None
    _ok(_lib_c().kc_tricount_direct(_p(_u32(nbr_u_direct)), _p(_u32(nbr_direct)), N, _p(tc),
                                    _p(None if left_son is None else _u32(left_son)),
                                    _p(None if right_son is None else _u32(right_son)), _p(tcd),
                                    tc_limit))
    return tc[:m], (None if tcd is None else tcd[: 2 * m])


def kmax_truss_GPU(ans, N, k, k_step, nbr_u_direct, nbr_direct, k_max=0, k_max_num=0):
    km, kn = _U(k_max), _U(k_max_num)
    _ok(_lib_c().kc_kmax_truss_GPU(ctypes.byref(km), ctypes.byref(kn), ans, N, k, k_step,
                                   _p(_u32(nbr_u_direct)), _p(_u32(nbr_direct))))
    return km.value, kn.value


def kmax_truss(N, nbr_u, nbr, deg_origin, launch_start, ans):
    """One pass of the reference's driver (src/kc_gpu.cpp:68-208); returns
    (k_max, k_max_num)."""
    node_delete = np.zeros(max(N, 1), dtype=np.uint32)
    k_min = kmax_core_gpu(N, nbr_u, nbr, deg_origin, node_delete, launch_start)
    deg_with_mask = np.zeros(max(N, 1), dtype=np.uint32)
    to = np.zeros(max(N, 1), dtype=np.uint32)
    li, new_N, deg, deg_direct = count_deg_with_mask(nbr, nbr_u, N, node_delete[:N],
                                                     deg_with_mask, to)
    nbr_u_direct = np.zeros(new_N + 1, dtype=np.uint32)
    nbr_u_direct[1:] = np.cumsum(deg_direct, dtype=np.uint64).astype(np.uint32)
    nbr_direct = build_nbr_direct(nbr, nbr_u, N, node_delete[:N], to[:N], nbr_u_direct, new_N)
    k, k_step = 0, K_MIDDLE
    if launch_start != NONE and k_min != 0:
        k_step = 1 if k_min >= 100 else K_MIDDLE - 2
        k = (k_min << 1) - (k_step * 2)
    if launch_start == NONE and k_min != 0:
        k_step = max(k_step, int((k_min - 1) / 2.1))
        if k_step < 200:
            k_step = max(K_MIDDLE, k_step // 2)
    return kmax_truss_GPU(ans, new_N, k, k_step, nbr_u_direct, nbr_direct)


def solve(N, nbr_u, nbr, deg_origin):
    """The reference's two passes (src/kc_gpu.cpp:336-341): (kmax, edges)."""
    k1, a1 = kmax_truss(N, nbr_u, nbr, deg_origin, NONE, 0)
    return kmax_truss(N, nbr_u, nbr, deg_origin, 0 if k1 == 0 else k1 - 1, a1)
