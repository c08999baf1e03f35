"""Device-side ingest of the edge-list text: the lines parsed in HBM equal the
host parser's and the oracle's (the reference's input contract,
src/kc_gpu.cpp:243-312) on regular files, odd separators, blank lines, CRLF,
extra columns, no final newline and files larger than one staging buffer; a
file with a single-number line goes to the host parser; the solve through
kt_solve_tsv equals the oracle's."""
import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _same_as_host(ctx, path, expect_device=True):
    from kc_cuda_amd import parse_tsv
    xs, ys, n, dev = ctx.parse_tsv_device(path)
    hxs, hys, hn = parse_tsv(path)
    assert dev == expect_device
    np.testing.assert_array_equal(xs, hxs)
    np.testing.assert_array_equal(ys, hys)
    assert n == hn
    oxs, oys = O.parse_tsv_bytes(path.read_bytes())
    np.testing.assert_array_equal(xs, oxs)
    np.testing.assert_array_equal(ys, oys)
    return xs, ys, n


def test_ingest_golden_files(ctx):
    import json
    expected = json.loads((GOLDEN / "tiny" / "expected.json").read_text())
    for name, exp in sorted(expected.items()):
        path = GOLDEN / "tiny" / name
        if not path.exists():
            continue
        print(f"[gpu] ingest {name}", flush=True)
        _same_as_host(ctx, path)
        r, _ = ctx.solve_tsv(path)
        assert f"kmax = {r.kmax}, Edges in kmax-truss = {r.truss_edges}." == exp, name


def test_ingest_odd_text(ctx, tmp_path):
    cases = {
        "crlf_extra": b"1\t0\r\n2,0\n0\t1\textra 9\n2\t1\n0\t2\n1\t2",
        "blank_lines": b"\n\n1 0\n \t\r\n2 0\n\n0 1\n2 1\n0 2\n1 2\n\n",
        "leading_seps": b"  1\t0\n\t2\t0\n,0 1\n2 1\n0 2\n1 2\n",
        "letters_as_digits": b"1 0\n2 0\n0 1\n2 1\n0 2\n1 2\n3a 0\n",
        "high_bytes": b"1 0\n2\xc3\xa90\n0 1\n2 1\n0 2\n1 2\n",
        "blank_only": b" \n\t\n",
    }
    for name, text in cases.items():
        path = tmp_path / f"{name}.tsv"
        path.write_bytes(text)
        print(f"[gpu] ingest {name}", flush=True)
        _same_as_host(ctx, path)


def test_ingest_single_number_line_goes_to_host(ctx, tmp_path):
    path = tmp_path / "single.tsv"
    path.write_bytes(b"1 0\n2\n0\n0 1\n2 1\n0 2\n1 2\n")
    xs, ys, n = _same_as_host(ctx, path, expect_device=False)
    r, _ = ctx.solve_tsv(path)
    off, adj, on = O.csr_from_lines(xs, ys)
    s = O.solve_csr(off, adj, on)
    assert (r.kmax, r.truss_edges) == (s.kmax, s.truss_edges)


def test_ingest_errors(ctx, tmp_path):
    from kc_cuda_amd import KtError
    empty = tmp_path / "empty.tsv"
    empty.write_bytes(b"")
    with pytest.raises(KtError, match="empty"):
        ctx.parse_tsv_device(empty)
    with pytest.raises(KtError, match="cannot open"):
        ctx.parse_tsv_device(tmp_path / "missing.tsv")
    bad = tmp_path / "bad.tsv"
    bad.write_bytes(b"0 1\n1 0\n9 1\n")  # x = 9 widens n; the graph is still solvable
    _same_as_host(ctx, bad)


def test_ingest_larger_than_staging(ctx, tmp_path):
    """RMAT-19 written as text (about 240 MB, several 64 MiB staging
    buffers): same lines as the host parser, same answer as the pairs."""
    from kc_cuda_amd import gen
    pairs, n = gen.rmat(19, 16, 0.57, 0.19, 0.19, True, 30, 5)
    path = tmp_path / "r19.tsv"
    nbytes = gen.write_tsv(path, pairs)
    print(f"[gpu] ingest rmat19 {nbytes} bytes", flush=True)
    assert nbytes > 3 * (64 << 20)
    from kc_cuda_amd import parse_tsv
    xs, ys, nn, dev = ctx.parse_tsv_device(path)
    hxs, hys, hn = parse_tsv(path)
    assert dev
    np.testing.assert_array_equal(xs, hxs)
    np.testing.assert_array_equal(ys, hys)
    assert nn == hn
    r, ms = ctx.solve_tsv(path)
    ref = ctx.solve_pairs(pairs, n)
    assert (r.kmax, r.truss_edges) == (ref.kmax, ref.truss_edges)
    print(f"[gpu] rmat19 text ingest {ms:.1f} ms", flush=True)
