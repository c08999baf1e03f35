"""CPU tests: workload generators, the TSV contract, the host TSV parser of
libkc_amd.so, and that the C-ABI library loads and exports every symbol its
headers declare (no device calls)."""
import ctypes
import json
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from oracle import oracle as O


def test_generators_deterministic_and_simple():
    from kc_cuda_amd import gen
    for f in (lambda s: gen.er_clique(2000, 8.0, 12, s), lambda s: gen.rmat(11, 8, seed=s),
              lambda s: gen.chung_lu(5000, 2.1, 30000, 10, s)):
        a, n = f(1)
        b, _ = f(1)
        c, _ = f(2)
        np.testing.assert_array_equal(a, b)
        assert not np.array_equal(a, c)
        assert a.dtype == np.uint32 and a.shape[1] == 2
        assert np.all(a[:, 0] < a[:, 1]) and np.all(a[:, 1] < n)
        keys = a[:, 0].astype(np.uint64) << np.uint64(32) | a[:, 1].astype(np.uint64)
        assert np.all(np.diff(keys.astype(np.int64) if keys.max() < 2**63 else keys) > 0)


def test_batch_generator_layout_and_cliques():
    """Batches: per-graph canonical sorted pairs, n in range, deterministic;
    every graph's kmax (oracle) is at least its planted clique's size floor."""
    from kc_cuda_amd import gen
    p, offs, ns = gen.batch(40, 16, 2000, 4.0, 20.0, 5, 12, 9)
    p2, offs2, ns2 = gen.batch(40, 16, 2000, 4.0, 20.0, 5, 12, 9)
    np.testing.assert_array_equal(p, p2)
    np.testing.assert_array_equal(offs, offs2)
    assert offs.shape == (41,) and offs[0] == 0 and offs[-1] == len(p)
    assert np.all((ns >= 16) & (ns <= 2000))
    for g in range(40):
        q = p[offs[g]:offs[g + 1]]
        assert np.all(q[:, 0] < q[:, 1]) and np.all(q[:, 1] < ns[g])
        keys = q[:, 0].astype(np.int64) << 32 | q[:, 1].astype(np.int64)
        assert np.all(np.diff(keys) > 0)
        if g < 8:
            assert O.solve_pairs(q, int(ns[g])).kmax >= min(5, int(ns[g]))


def test_serve_mix_sizes_and_layout():
    """Serving queries: deterministic, canonical pairs, lognormal sizes near
    the requested median, planted cliques that bound kmax from below."""
    from kc_cuda_amd import gen
    a = gen.serve_mix(12, 20_000, 1.0, 200_000, 5)
    b = gen.serve_mix(12, 20_000, 1.0, 200_000, 5)
    for (p, n), (q, m) in zip(a, b):
        np.testing.assert_array_equal(p, q)
        assert n == m and p.dtype == np.uint32
        assert np.all(p[:, 0] < p[:, 1]) and np.all(p[:, 1] < n)
    sizes = np.array([len(p) for p, _ in a])
    assert 5_000 < np.median(sizes) < 80_000 and sizes.max() <= 300_000
    p, n = gen.serve_query(5_000, "chung_lu", 12, 3)
    assert O.solve_pairs(p, n).kmax >= 12


def test_er_density_and_clique():
    from kc_cuda_amd import gen
    p, n = gen.er_clique(16384, 16.0, 24, 0)
    assert abs(len(p) - 16384 * 8) < 0.02 * 16384 * 8 + 300
    s = O.solve_pairs(p, n)
    assert s.kmax >= 24


def test_rmat_matches_recorded_sizes():
    from kc_cuda_amd import gen
    g = json.loads((GOLDEN / "kc_cpu_results.json").read_text())
    p, n = gen.rmat(12)
    assert len(p) == g["rmat12"]["m"] and n == 4096


def test_tsv_contract(tmp_path):
    from kc_cuda_amd import gen
    p, n = gen.er_clique(500, 6.0, 8, 4)
    path = tmp_path / "g.tsv"
    gen.write_tsv(path, p)
    lines = path.read_text().splitlines()
    assert len(lines) == 2 * len(p)
    xy = np.array([list(map(int, l.split("\t"))) for l in lines], dtype=np.int64)
    # grouped by the second column ascending, ties by the first
    order = np.lexsort((xy[:, 0], xy[:, 1]))
    np.testing.assert_array_equal(order, np.arange(len(xy)))
    got = {(min(a, b), max(a, b)) for a, b in xy}
    assert got == {tuple(map(int, e)) for e in p}


def test_host_parser_matches_oracle(tmp_path):
    from kc_cuda_amd import gen, parse_tsv
    p, n = gen.rmat(10, 8, seed=3)
    path = tmp_path / "r.tsv"
    gen.write_tsv(path, p)
    xs, ys, nn = parse_tsv(path)
    oxs, oys = O.parse_tsv_bytes(path.read_bytes())
    np.testing.assert_array_equal(xs, oxs)
    np.testing.assert_array_equal(ys, oys)
    odd = tmp_path / "odd.tsv"
    odd.write_bytes(b"1 0\r\n2,0\n0\t1\textra\n2\t1\n0\t2\n1\t2")
    xs, ys, nn = parse_tsv(odd)
    oxs, oys = O.parse_tsv_bytes(odd.read_bytes())
    assert xs.tolist() == oxs.tolist() and ys.tolist() == oys.tolist() and nn == 3


def test_library_exports_every_declared_symbol():
    import kc_cuda_amd
    lib = kc_cuda_amd.native_library()
    declared = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        if "extern \"C\"" not in text:
            continue
        declared |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b((?:kt|kc)_\w+)\s*\(", text, re.M))
    assert declared, "no C declarations found"
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(kc_cuda_amd.EXPORTS) <= declared


def test_cpp_abi_symbols_present():
    """The reference's kc_gpu.h functions are exported with C++ linkage."""
    so = ROOT / "kc_cuda_amd" / "lib" / "libkc_amd.so"
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", str(so)], capture_output=True,
                         text=True).stdout
    for fn in ("kmax_core_gpu(", "count_deg_with_mask(", "build_nbr_direct(",
               "build_graph_from_direct(", "tricount_direct(", "kmax_truss_GPU(", "initGPU(",
               "test_cuda("):
        if fn not in out:
            pytest.skip(f"{fn} not built yet")
        assert fn in out


def test_cli_usage_exit_code():
    from kc_cuda_amd._lib import kc_gpu_binary
    r = subprocess.run([str(kc_gpu_binary())], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stderr


def test_no_device_fails_loudly():
    import kc_cuda_amd
    lib = kc_cuda_amd.native_library()
    cnt = ctypes.c_int(0)
    if lib.kt_device_count(ctypes.byref(cnt)) == 0 and cnt.value > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(kc_cuda_amd.KtError):
        kc_cuda_amd.Context(0)


def test_missing_library_raises(monkeypatch, tmp_path):
    from kc_cuda_amd import _lib
    monkeypatch.setattr(_lib, "LIBDIR", tmp_path)
    monkeypatch.setattr(_lib, "_cache", {})
    with pytest.raises(_lib.NativeLibraryMissing):
        _lib.load_device()
