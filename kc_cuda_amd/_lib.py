"""Locate and load the in-tree native libraries.

The HIP library is never replaced by a Python or CPU fallback: when it is
missing, ``load_device`` raises with the build command to run.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIBDIR = PKG / "lib"
BINDIR = PKG / "bin"
DEVICE_LIB = "libkc_amd.so"

_cache: dict[str, ctypes.CDLL] = {}


class NativeLibraryMissing(RuntimeError):
    pass


def _load(name: str) -> ctypes.CDLL:
    if name in _cache:
        return _cache[name]
    path = LIBDIR / name
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            f"or `python -m kc_cuda_amd.build` from the repository root")
    lib = ctypes.CDLL(str(path), mode=os.RTLD_LOCAL)
    _cache[name] = lib
    return lib


def load_host(name: str) -> ctypes.CDLL:
    return _load(name)


def load_device() -> ctypes.CDLL:
    return _load(DEVICE_LIB)


def device_lib_path() -> Path:
    return LIBDIR / DEVICE_LIB


def kc_gpu_binary() -> Path:
    return BINDIR / "kc_gpu"
