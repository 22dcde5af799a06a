"""CSR model on the CPU: validate_csr, from_edge_pairs, from_edge_pairs_oriented and
parse_tsv_to_edges, ported from crates/grafeo-cuda-core/tests/csr_export.rs (the cases that do
not need a Grafeo store), plus property tests against a set model and against the oracle."""

from __future__ import annotations

import itertools
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from grafeo_cuda import _native
from grafeo_cuda._native import GcError
from graphs import csr_sha256

# ---- validate_csr (csr_export.rs:533-562) -------------------------------------


def invalid(offsets, targets) -> str:
    with pytest.raises(GcError) as info:
        _native.validate_csr(offsets, targets)
    assert info.value.kind == "InvalidCsr"
    return str(info.value)


def test_validate_accepts_well_formed_csr():
    _native.validate_csr([0], [])
    _native.validate_csr([0, 0, 0], [])
    _native.validate_csr([0, 3, 6, 9, 12], [1, 2, 3, 0, 2, 3, 0, 1, 3, 0, 1, 2])
    _native.validate_csr([0, 2, 3], [0, 1, 0])  # self-loops are allowed


def test_validate_rejects_malformed_csr():
    assert "empty" in invalid([], [])
    assert "offsets[0]" in invalid([1, 1], [0])
    assert "offsets[1] is 2" in invalid([0, 2], [1])
    assert "decrease" in invalid([0, 2, 1, 2], [1, 2])
    assert "only 2 nodes" in invalid([0, 1, 1], [2])
    assert "not strictly increasing" in invalid([0, 2, 2], [1, 0])
    assert "not strictly increasing" in invalid([0, 2, 2], [1, 1])


def test_validate_messages_match_the_oracle():
    rng = np.random.default_rng(3)
    for _ in range(300):
        n = int(rng.integers(0, 6))
        off = np.sort(rng.integers(0, 6, size=n + 1)) if rng.random() < 0.7 else rng.integers(0, 6, size=n + 1)
        if rng.random() < 0.5 and n + 1 > 0:
            off[0] = 0
        m = int(off[-1]) if n + 1 > 0 and rng.random() < 0.8 else int(rng.integers(0, 6))
        tgt = rng.integers(0, max(n, 1) + 1, size=m)
        expected = oracle.validate_csr(off, tgt)
        try:
            _native.validate_csr(off, tgt)
            got = None
        except GcError as e:
            got = str(e).removeprefix("invalid CSR: ")
        assert got == expected, (off, tgt)


# ---- from_edge_pairs (csr_export.rs:381-400, 463-474) ---------------------------


def set_model(edges, n, symmetric):
    sets = [set() for _ in range(n)]
    for s, d in edges:
        sets[s].add(d)
        if symmetric:
            sets[d].add(s)
    offsets = [0]
    targets = []
    for s in sets:
        targets += sorted(s)
        offsets.append(len(targets))
    return offsets, targets


def test_edge_pairs_keep_isolated_trailing_nodes():
    off, tgt = _native.from_edge_pairs([(0, 1)], 4, True)
    assert off.tolist() == [0, 1, 2, 2, 2] and tgt.tolist() == [1, 0]


def test_edge_pairs_empty():
    off, tgt = _native.from_edge_pairs([], 0, True)
    assert off.tolist() == [0] and tgt.tolist() == []
    off, tgt = _native.from_edge_pairs_oriented([], 0)
    assert off.tolist() == [0] and tgt.tolist() == []
    off, tgt = _native.from_edge_pairs_oriented([(2, 2)], 3)
    assert off.tolist() == [0, 0, 0, 0] and tgt.tolist() == []


def test_self_loops_appear_once_and_parallel_edges_collapse():
    off, tgt = _native.from_edge_pairs([(0, 0), (0, 1)], 2, True)
    assert (off.tolist(), tgt.tolist()) == ([0, 2, 3], [0, 1, 0])
    off, tgt = _native.from_edge_pairs([(0, 0), (0, 1)], 2, False)
    assert (off.tolist(), tgt.tolist()) == ([0, 2, 2], [0, 1])
    off, tgt = _native.from_edge_pairs([(0, 1), (0, 1), (1, 0)], 2, True)
    assert (off.tolist(), tgt.tolist()) == ([0, 1, 2], [1, 0])


def test_edge_pairs_reject_ids_out_of_range():
    with pytest.raises(GcError) as info:
        _native.from_edge_pairs([(0, 5)], 3, True)
    assert info.value.kind == "InvalidCsr"


def test_edge_pairs_agree_with_set_model_and_oracle():
    """edge_pairs_agree_with_store_export (csr_export.rs:381-400): the store export is the set
    model of csr_export.rs:334-360 for a store without deletions."""
    rng = np.random.default_rng(11)
    for _ in range(300):
        n = int(rng.integers(1, 12))
        edges = [tuple(map(int, p)) for p in rng.integers(0, n, size=(int(rng.integers(0, 40)), 2))]
        for symmetric in (False, True):
            off, tgt = _native.from_edge_pairs(edges, n, symmetric)
            _native.validate_csr(off, tgt)
            e_off, e_tgt = set_model(edges, n, symmetric)
            assert off.tolist() == e_off and tgt.tolist() == e_tgt
            o_off, o_tgt = oracle.from_edge_pairs(edges, n, symmetric)
            assert off.tolist() == o_off.tolist() and tgt.tolist() == o_tgt.tolist()


def test_edge_pairs_large_input_uses_the_same_order():
    """Inputs above the radix-sort threshold give the same CSR as the oracle's comparison sort."""
    pairs = _native.gen_rmat(13, 16, 5)
    off, tgt = _native.from_edge_pairs(pairs, 1 << 13, True)
    o_off, o_tgt = oracle.from_edge_pairs(pairs, 1 << 13, True)
    assert np.array_equal(off, o_off) and np.array_equal(tgt, o_tgt)
    off, tgt = _native.from_edge_pairs_oriented(pairs, 1 << 13)
    o_off, o_tgt = oracle.from_edge_pairs_oriented(pairs, 1 << 13)
    assert np.array_equal(off, o_off) and np.array_equal(tgt, o_tgt)


def test_oriented_csr_stores_each_edge_once_and_counts_triangles():
    """csr_export.rs:402-458."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = int(rng.integers(1, 14))
        edges = [tuple(map(int, p)) for p in rng.integers(0, n, size=(int(rng.integers(0, 60)), 2))]
        off, tgt = _native.from_edge_pairs_oriented(edges, n)
        _native.validate_csr(off, tgt)
        undirected = {(min(s, d), max(s, d)) for s, d in edges if s != d}
        assert len(tgt) == len(undirected)
        degree = [0] * n
        for u, v in undirected:
            degree[u] += 1
            degree[v] += 1
        seen = set()
        for u in range(n):
            for v in tgt[off[u]:off[u + 1]]:
                v = int(v)
                assert degree[u] < degree[v] or (degree[u] == degree[v] and u < v)
                key = (min(u, v), max(u, v))
                assert key not in seen
                seen.add(key)
        assert seen == undirected
        brute = sum(1 for a, b, c in itertools.combinations(range(n), 3)
                    if (a, b) in undirected and (b, c) in undirected and (a, c) in undirected)
        assert oracle.triangle_count_oriented(off, tgt) == brute
        o_off, o_tgt = oracle.from_edge_pairs_oriented(edges, n)
        assert off.tolist() == o_off.tolist() and tgt.tolist() == o_tgt.tolist()


# ---- parse_tsv_to_edges (csr_export.rs:495-531) ---------------------------------


def test_tsv_parser_remaps_ids_in_order_of_appearance(tmp_path):
    p = tmp_path / "remap.tsv"
    p.write_text("# comment\n% matrix-market style comment\n\n19\t88\t1.0\n88 3\nnot numbers\n19\n3\t19\n")
    edges, nodes = _native.parse_tsv_to_edges(str(p))
    assert nodes == 3
    assert edges.tolist() == [[0, 1], [1, 2], [2, 0]]
    assert oracle.parse_tsv_to_edges(str(p)) == ([(0, 1), (1, 2), (2, 0)], 3)


def test_tsv_parser_handles_large_ids_and_empty_files(tmp_path):
    p = tmp_path / "large.tsv"
    p.write_text("18446744073709551615 0\n")
    edges, nodes = _native.parse_tsv_to_edges(str(p))
    assert (edges.tolist(), nodes) == ([[0, 1]], 2)
    p = tmp_path / "empty.tsv"
    p.write_text("# nothing here\n")
    edges, nodes = _native.parse_tsv_to_edges(str(p))
    assert (edges.tolist(), nodes) == ([], 0)


def test_tsv_parser_reports_missing_file(tmp_path):
    with pytest.raises(GcError) as info:
        _native.parse_tsv_to_edges(str(tmp_path / "definitely-missing.tsv"))
    assert info.value.kind == "Export"


def test_tsv_parser_matches_the_python_restatement(tmp_path):
    rng = np.random.default_rng(2)
    tokens = ["1", "2", "+3", "-4", "x", "18446744073709551616", "007", "5.0", "", "#", "%"]
    lines = []
    for _ in range(400):
        k = int(rng.integers(0, 4))
        sep = [" ", "\t", "  "][int(rng.integers(0, 3))]
        lines.append(sep.join(tokens[int(i)] for i in rng.integers(0, len(tokens), size=k)))
    p = tmp_path / "fuzz.tsv"
    p.write_text("\r\n".join(lines))
    edges, nodes = _native.parse_tsv_to_edges(str(p))
    o_edges, o_nodes = oracle.parse_tsv_to_edges(str(p))
    assert [tuple(e) for e in edges.tolist()] == o_edges and nodes == o_nodes


# Rust's parse on text with Unicode: fields separated and lines trimmed by the White_Space
# code points (U+0085, U+00A0, U+1680, U+2000..200A, U+2028, U+2029, U+202F, U+205F, U+3000
# besides ASCII's), not by U+001C..001F or U+200B (zero-width space, not White_Space), which
# then belong to the field and make it fail to parse.
UNICODE_LINES = [
    ("1\u00a02", (1, 2)),
    ("2\u20033", (2, 3)),
    ("\u30003\u3000\u30004\u3000", (3, 4)),
    ("4\u16805\u0085", (4, 5)),
    ("5\u20286", (5, 6)),
    ("6\u202f\u205f7", (6, 7)),
    ("7\x1c8", None),
    ("8\u200b9", None),
    ("\u00a0# comment after a no-break space", None),
    ("9 \u00e910", None),
    ("10\u200a11", (10, 11)),
]


def unicode_file(tmp_path, name="unicode.tsv"):
    p = tmp_path / name
    p.write_bytes("\n".join(line for line, _ in UNICODE_LINES).encode("utf-8"))
    ids, want = {}, []
    for _, pair in UNICODE_LINES:
        if pair:
            want.append((ids.setdefault(pair[0], len(ids)), ids.setdefault(pair[1], len(ids))))
    return p, want, len(ids)


def test_tsv_parser_unicode_whitespace(tmp_path):
    p, want, nodes = unicode_file(tmp_path)
    edges, n = _native.parse_tsv_to_edges(str(p))
    assert [tuple(e) for e in edges.tolist()] == want and n == nodes
    assert oracle.parse_tsv_to_edges(str(p)) == (want, nodes)


@pytest.mark.parametrize("bad", [b"\xff", b"\xc0\x80", b"\xed\xa0\x80", b"\xf4\x90\x80\x80", b"\xe2\x80",
                                 b"\x80", b"\xf0\x8f\xbf\xbf"])
def test_tsv_parser_rejects_invalid_utf8(tmp_path, bad):
    """Rust's BufRead::lines fails on a line that is not UTF-8 (a lone continuation byte, an
    overlong form, a surrogate, a code point above U+10FFFF, a truncated sequence), even in a
    comment line, and the whole parse fails with it."""
    p = tmp_path / "bad.tsv"
    p.write_bytes(b"1 2\n# comment " + bad + b"\n3 4\n")
    with pytest.raises(GcError) as info:
        _native.parse_tsv_to_edges(str(p))
    assert info.value.kind == "Export" and "valid UTF-8" in str(info.value)
    with pytest.raises(UnicodeDecodeError):
        oracle.parse_tsv_to_edges(str(p))


def test_tsv_parser_accepts_every_valid_utf8_form(tmp_path):
    p = tmp_path / "ok.tsv"
    p.write_bytes("# \u00e9 \u20ac \U0001f600 \U0010ffff \ud7ff \ue000\n1 2\n".encode("utf-8"))
    edges, n = _native.parse_tsv_to_edges(str(p))
    assert (edges.tolist(), n) == ([[0, 1]], 2)


# ---- generators ------------------------------------------------------------------


def test_generators_are_deterministic_and_thread_independent():
    a = _native.gen_rmat(12, 8, 42)
    b = _native.gen_rmat(12, 8, 42)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, _native.gen_rmat(12, 8, 43))
    assert (a[:, 0] != a[:, 1]).all()  # no self-loops
    code = (
        "import sys; sys.path.insert(0, 'grafeo-cuda_amd'); from grafeo_cuda import _native; "
        "import hashlib; print(hashlib.sha256(_native.gen_rmat(12, 8, 42).tobytes()).hexdigest())"
    )
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    one = subprocess.run([sys.executable, "-c", code], cwd=repo, capture_output=True, text=True,
                         env={**os.environ, "GC_HOST_THREADS": "1"}, check=True).stdout.strip()
    import hashlib

    assert one == hashlib.sha256(a.tobytes()).hexdigest()


def test_rmat_quadrant_frequencies():
    """At each bit level, (src bit, dst bit) = (0,0), (0,1), (1,0), (1,1) with probabilities
    (a, b, c, d) = (0.57, 0.19, 0.19, 0.05), before the id permutation."""
    pairs = _native.gen_rmat(16, 16, 9, permute=False).astype(np.int64)
    for level in (0, 7, 15):
        s = (pairs[:, 0] >> level) & 1
        d = (pairs[:, 1] >> level) & 1
        freq = [np.mean((s == i) & (d == j)) for i, j in ((0, 0), (0, 1), (1, 0), (1, 1))]
        # self-loops were dropped, which skews the estimate by far less than this tolerance
        assert np.allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.01), (level, freq)


def test_grid_generator():
    pairs = _native.gen_grid(3, 4)
    assert len(pairs) == 3 * 3 + 2 * 4
    off, tgt = _native.from_edge_pairs(pairs, 12, True)
    assert np.diff(off).tolist() == [2, 3, 3, 2, 3, 4, 4, 3, 2, 3, 3, 2]


def test_generated_inputs_match_the_golden_fixtures():
    import json

    from graphs import er_graph, rmat_graph

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    er = json.load(open(os.path.join(here, "er4096.json")))
    assert csr_sha256(*er_graph(4096, 16, seed=1)) == er["csr_sha256"]
    rm = json.load(open(os.path.join(here, "rmat_counts.json")))
    for scale in (10, 12, 14):
        assert csr_sha256(*rmat_graph(scale, 16, seed=1)) == rm[f"rmat{scale}_ef16_seed1"]["csr_sha256"]
