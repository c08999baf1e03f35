"""Build the native parts in-tree (no JIT cache, so the outputs travel).

    python -m kc_cuda_amd.build            # everything
    python -m kc_cuda_amd.build --quiet

Outputs:
    kc_cuda_amd/lib/libkc_amd.so   HIP kernels (gfx950) + host orchestration + C ABI
    kc_cuda_amd/lib/libktgen.so    synthetic graph generators (host, OpenMP)
    kc_cuda_amd/bin/kc_gpu         `kc_gpu -f <file>` command line
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
BIN = PKG / "bin"
OBJ = PKG / "_build"
ROOT = PKG.parent
ARCH = os.environ.get("KT_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -packed-fp32-ops: keep the compiler from emitting packed fp32 VALU ops.
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Wall",
             "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             f"-I{ROOT / 'include'}", f"-I{CSRC}"]

DEVICE_SOURCES = ["kt_kernels.hip", "kt_peel.hip", "kt_ingest.hip", "kt_batch_union.hip", "kt_solver.hip", "kt_stages.hip",
                  "kt_compat.hip", "kt_noalloc.hip", "kt_batch_host.cpp", "kt_serve.cpp", "kt_capi.cpp", "kt_tsv.cpp"]


def _run(cmd: list[str], quiet: bool) -> None:
    if not quiet:
        print("+", " ".join(str(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({r.returncode}): {' '.join(map(str, cmd))}\n{out}")
    if out.strip() and not quiet:
        print(out)


def _compile(src: Path, quiet: bool) -> Path:
    obj = OBJ / (src.name + ".o")
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    _run([HIPCC, *HIP_FLAGS, *lang, "-c", str(src), "-o", str(obj)], quiet)
    return obj


def build_device(quiet: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    BIN.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in DEVICE_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, quiet), srcs))
    so = LIB / "libkc_amd.so"
    tmp = LIB / "libkc_amd.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-lpthread"], quiet)
    os.replace(tmp, so)
    cli = BIN / "kc_gpu"
    _run([HIPCC, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", "-o", str(cli),
          str(CSRC / "kc_gpu_main.cpp"), f"-L{LIB}", "-lkc_amd", "-Wl,-rpath,$ORIGIN/../lib"], quiet)
    return so


def build_gen(quiet: bool = False) -> Path:
    LIB.mkdir(exist_ok=True)
    so = LIB / "libktgen.so"
    tmp = LIB / "libktgen.so.tmp"
    _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-Wall", "-o", str(tmp),
          str(CSRC / "ktgen.cpp")], quiet)
    os.replace(tmp, so)
    return so


def build_all(quiet: bool = False) -> None:
    build_gen(quiet)
    build_device(quiet)


if __name__ == "__main__":
    build_all(quiet="--quiet" in sys.argv)
