"""TSV edge lists parsed on the GPU (gc_parse_tsv_to_edges_device, gc_graph_from_tsv): the same
pairs and node count as the host parser (parse_tsv_to_edges,
crates/grafeo-cuda-core/src/csr_export.rs:246-287, itself pinned by the reference's own parse
cases in test_csr.py) on edge cases, random files and a full-size file, and the same error for a
missing file; the same White_Space separators and the same error for a file that is not
UTF-8, as Rust's lines, trim and split_whitespace give the reference."""

from __future__ import annotations

import numpy as np
import pytest

import grafeo_cuda
from grafeo_cuda import _native

pytestmark = pytest.mark.gpu


def same(device, path):
    host = _native.parse_tsv_to_edges(str(path))
    dev = _native.parse_tsv_to_edges_device(device._handle, str(path))
    assert host[1] == dev[1], "node count"
    assert np.array_equal(host[0], dev[0]), "pairs"
    return dev


EDGE_CASES = (
    "# a comment\n"
    "% another comment\n"
    "\n"
    "   \t  \n"
    "10 20\n"
    "20\t30\r\n"
    "  30   10  \n"
    "+40 +10\n"
    "a 1\n"
    "1 b\n"
    "-1 2\n"
    "7\n"
    "18446744073709551615 5\n"
    "18446744073709551616 5\n"
    "5 18446744073709551615 99 extra\n"
    "20 20\n"
    "#10 20\n"
    "10 20 # trailing words\n"
    "\v10\f 40\n"
    "0 0"
)


def test_edge_cases(device, tmp_path):
    path = tmp_path / "cases.tsv"
    path.write_bytes(EDGE_CASES.encode())
    pairs, n = same(device, path)
    # ids in order of first appearance: 10, 20, 30, 40, 18446744073709551615, 5, 0
    assert n == 7
    assert pairs.tolist() == [[0, 1], [1, 2], [2, 0], [3, 0], [4, 5], [5, 4], [1, 1], [0, 1], [0, 3], [6, 6]]


def test_empty_and_comment_only_files(device, tmp_path):
    for i, text in enumerate(("", "\n", "# only\n% comments\n", "   ", "1 2", "1 2\n")):
        path = tmp_path / f"f{i}.tsv"
        path.write_bytes(text.encode())
        same(device, path)


def test_random_files(device, tmp_path):
    rng = np.random.default_rng(21)
    # the last four span several 64-byte chunks of the file: a long comment, fields far apart,
    # ids with many leading zeros (valid for u64::from_str) and a long field that overflows
    kinds = ["{a}\t{b}", "{a} {b}", "  {a}  {b}  ", "{a} {b} {c}", "# {a}", "%", "", "x {b}", "{a}", "+{a} {b}",
             "{a}\t{b}\r", "#" + "x" * 300, "{a}" + " " * 150 + "{b}", "0" * 90 + "{a} 000{b}", "9" * 70 + " {b}"]
    for t in range(12):
        lines = []
        ids = rng.integers(0, [50, 10**6, 2**40][t % 3], size=3000)
        for k in range(3000):
            kind = kinds[int(rng.integers(0, len(kinds)))]
            lines.append(kind.format(a=ids[k], b=ids[(k * 7 + 3) % 3000], c=ids[(k * 13) % 3000]))
        path = tmp_path / f"r{t}.tsv"
        path.write_bytes(("\n".join(lines) + ("\n" if t % 2 else "")).encode())
        same(device, path)


def test_full_size_file_and_graph(device, tmp_path):
    """An R-MAT 18 edge list with sparse 64-bit ids (about 2 M lines): the same pairs, and the
    graph built on the GPU from the file counts the same triangles as the host path."""
    pairs = _native.gen_rmat(18, 8, 3).astype(np.uint64)
    rng = np.random.default_rng(22)
    remap = rng.integers(0, 2**62, size=1 << 18, dtype=np.uint64)
    path = tmp_path / "rmat18.tsv"
    with open(path, "w") as f:
        f.write("# R-MAT 18\n")
        np.savetxt(f, remap[pairs], fmt="%d", delimiter="\t")
    host_pairs, n = same(device, path)
    off, tgt = _native.from_edge_pairs(host_pairs, n, True)
    expected = grafeo_cuda.GpuGraph.from_csr(off, tgt, device).triangle_count()["total_triangles"]
    g = grafeo_cuda.GpuGraph.from_tsv(str(path), device)
    assert g.parsed_pairs == len(host_pairs) and g.num_nodes == n and g.num_edges == len(tgt)
    assert g.triangle_count()["total_triangles"] == expected


def test_missing_file_is_an_export_error(device, tmp_path):
    path = str(tmp_path / "missing.tsv")
    with pytest.raises(_native.GcError) as host_err:
        _native.parse_tsv_to_edges(path)
    with pytest.raises(_native.GcError) as dev_err:
        _native.parse_tsv_to_edges_device(device._handle, path)
    assert host_err.value.kind == dev_err.value.kind == "Export"


UNICODE_SEPARATORS = ["\u00a0", "\u2003", "\u3000", "\u1680", "\u2000", "\u200a", "\u2028", "\u205f",
                      "\u0085", " \u202f\t"]
NOT_SEPARATORS = ["\x1c", "\u200b", "\u00e9"]


def test_unicode_whitespace(device, tmp_path):
    """White_Space code points separate fields and are trimmed (Rust's trim and
    split_whitespace), in random lines with every separator, some spanning 64-byte chunks;
    other code points stay inside their field, which then fails to parse. The device parser
    gives the host parser's pairs, which the CPU tests check against the restatement."""
    rng = np.random.default_rng(31)
    for t in range(6):
        lines = []
        for k in range(2000):
            a, b = (int(x) for x in rng.integers(0, 400, size=2))
            sep = UNICODE_SEPARATORS[int(rng.integers(0, len(UNICODE_SEPARATORS)))]
            lead = UNICODE_SEPARATORS[int(rng.integers(0, len(UNICODE_SEPARATORS)))] * int(rng.integers(0, 3))
            if k % 17 == 0:
                sep = NOT_SEPARATORS[k % len(NOT_SEPARATORS)]
            if k % 101 == 0:
                sep = sep * 40
            lines.append(f"{lead}{a}{sep}{b}{lead}")
        path = tmp_path / f"u{t}.tsv"
        path.write_bytes("\n".join(lines).encode("utf-8"))
        pairs, n = same(device, path)
        assert len(pairs) > 1000


@pytest.mark.parametrize("bad", [b"\xff", b"\xc0\x80", b"\xed\xa0\x80", b"\xf4\x90\x80\x80", b"\xe2\x80", b"\x80"])
def test_invalid_utf8_is_an_export_error(device, tmp_path, bad):
    path = tmp_path / "bad.tsv"
    path.write_bytes(b"1 2\n" + b"x" * 100 + b" # " + bad + b"\n3 4\n")
    with pytest.raises(_native.GcError) as host_err:
        _native.parse_tsv_to_edges(str(path))
    with pytest.raises(_native.GcError) as dev_err:
        _native.parse_tsv_to_edges_device(device._handle, str(path))
    assert host_err.value.kind == dev_err.value.kind == "Export"
    assert str(host_err.value) == str(dev_err.value)
    with pytest.raises(RuntimeError, match="valid UTF-8"):  # the Python surface's error type
        grafeo_cuda.GpuGraph.from_tsv(str(path), device)
