"""TSV edge lists parsed on the GPU (gc_parse_tsv_to_edges_device, gc_graph_from_tsv): the same
pairs and node count as the host parser (parse_tsv_to_edges,
crates/grafeo-cuda-core/src/csr_export.rs:246-287, itself pinned by the reference's own parse
cases in test_csr.py) on edge cases, random files and a full-size file, and the same error for a
missing file."""

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
