"""Pins of the CPU oracle (oracle/oracle.c), so that it can serve as the checker of the GPU path.

Every oracle function is checked against something it was not derived from:
  - the reference's own known answers (tests/roundtrip.rs, correctness.rs:330-386),
  - closed forms (C(n,3) triangles of K_n, K_n is an n-truss, path/grid BFS distances,
    PageRank 1/n on regular graphs, the star's 2x2 linear system),
  - brute force on small random graphs (O(n^3) triangles, peeling by definition),
  - networkx / scipy results written to tests/golden/ by tests/golden/make_golden.py,
  - and each *_fast variant against its literal *_ref twin.
"""

from __future__ import annotations

import itertools
import json
import math
import os

import numpy as np
import pytest

import oracle
from graphs import catalog, complete_edges, csr_sha256, er_graph, grid_graph, path_edges, random_graph
from graphs import rmat_graph, undirected

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def brute_triangles(off, tgt):
    n = len(off) - 1
    adj = [set(int(v) for v in tgt[off[u]:off[u + 1]]) for u in range(n)]
    return sum(1 for a, b, c in itertools.combinations(range(n), 3) if b in adj[a] and c in adj[a] and c in adj[b])


def brute_truss(off, tgt, k):
    """Peel by definition: repeatedly drop every edge in fewer than k-2 triangles of the
    remaining loop-free graph; the fixpoint is the maximal k-truss."""
    n = len(off) - 1
    edges = {(u, int(v)) for u in range(n) for v in tgt[off[u]:off[u + 1]] if int(v) > u}
    while True:
        adj = {u: set() for u in range(n)}
        for u, v in edges:
            adj[u].add(v)
            adj[v].add(u)
        weak = {(u, v) for u, v in edges if len(adj[u] & adj[v]) < k - 2}
        if not weak:
            return sorted(edges)
        edges -= weak


# ---- triangle counting ------------------------------------------------------------


CLOSED_FORM_TRIANGLES = {
    "single node": 0, "edgeless": 0, "K4": 4, "K5": 10, "K7": 35, "path": 0, "petersen": 0,
    "star": 0, "wheel": 6, "diamond": 2, "bowtie": 2, "triangle, edge and isolated nodes": 1,
    "two K4 joined by a bridge": 8, "K4 with a tail": 4, "messy": 3,
}


def test_triangles_on_catalog_closed_forms_and_networkx():
    golden = {g["name"]: g for g in load("catalog.json")}
    for name, off, tgt in catalog():
        assert golden[name]["csr_sha256"] == csr_sha256(off, tgt)
        expected = CLOSED_FORM_TRIANGLES[name]
        assert golden[name]["triangles"] == expected, name
        assert oracle.triangle_count(off, tgt) == expected, name
        assert oracle.triangle_count(off, tgt, fast=True) == expected, name


def test_triangles_of_complete_graphs():
    for n in range(1, 30, 4):
        off, tgt = undirected(n, complete_edges(n))
        assert oracle.triangle_count(off, tgt) == math.comb(n, 3)
        assert oracle.triangle_count(off, tgt, fast=True) == math.comb(n, 3)
        assert oracle.triangle_count_oriented(off, tgt) == 6 * math.comb(n, 3)


def test_triangles_brute_force_on_random_graphs():
    rng = np.random.default_rng(1)
    for _ in range(300):
        n, edges = random_graph(rng, max_nodes=14, max_edges=60)
        off, tgt = undirected(n, edges)
        b = brute_triangles(off, tgt)
        assert oracle.triangle_count(off, tgt) == b
        assert oracle.triangle_count(off, tgt, fast=True) == b
        # symmetric CSR: every triangle appears 6 times as (entry, common neighbour), loops add
        # |N(u)| - 1 for each loop entry u->u plus matches through loop entries
        loops = [u for u in range(n) if u in set(int(v) for v in tgt[off[u]:off[u + 1]])]
        if not loops:
            assert oracle.triangle_count_oriented(off, tgt) == 6 * b


def test_triangle_count_directed_semantics_brute_force():
    """On a non-symmetric CSR, triangle_count counts (u<v<w) with v, w in N(u), w in N(v)."""
    from grafeo_cuda import _native

    rng = np.random.default_rng(4)
    for _ in range(200):
        n, edges = random_graph(rng, max_nodes=12, max_edges=50)
        off, tgt = _native.from_edge_pairs(edges, n, False)
        adj = [set(int(v) for v in tgt[off[u]:off[u + 1]]) for u in range(n)]
        brute = sum(1 for u, v, w in itertools.combinations(range(n), 3) if v in adj[u] and w in adj[u] and w in adj[v])
        assert oracle.triangle_count(off, tgt) == brute
        oriented = sum(len(adj[u] & adj[v]) for u in range(n) for v in adj[u])
        assert oracle.triangle_count_oriented(off, tgt) == oriented


def test_triangles_against_networkx_fixtures():
    er = load("er4096.json")
    off, tgt = er_graph(4096, 16, seed=1)
    assert csr_sha256(off, tgt) == er["csr_sha256"]
    assert oracle.triangle_count(off, tgt) == er["triangles"]
    assert oracle.triangle_count(off, tgt, fast=True) == er["triangles"]
    rm = load("rmat_counts.json")
    for scale in (10, 12, 14):
        e = rm[f"rmat{scale}_ef16_seed1"]
        assert e["source"] == "networkx"
        off, tgt = rmat_graph(scale, 16, seed=1)
        assert oracle.triangle_count(off, tgt) == e["triangles"]
        assert oracle.triangle_count(off, tgt, fast=True) == e["triangles"]


# ---- BFS ----------------------------------------------------------------------------


def test_bfs_known_answers():
    off, tgt = undirected(4, complete_edges(4))
    dist, reach, depth = oracle.bfs(off, tgt, 0)
    assert dist.tolist() == [0, 1, 1, 1] and (reach, depth) == (4, 1)
    off, tgt = undirected(10, path_edges(10))
    for s in range(10):
        dist, reach, depth = oracle.bfs(off, tgt, s)
        assert dist.tolist() == [abs(i - s) for i in range(10)]
        assert depth == max(s, 9 - s) and reach == 10
    dist, reach, depth = oracle.bfs([0, 0, 0, 0], [], 1)
    assert dist.tolist() == [-1, 0, -1] and (reach, depth) == (1, 0)
    with pytest.raises(ValueError):
        oracle.bfs([0], [], 0)


def test_bfs_grid_is_manhattan_distance():
    rows, cols = 37, 23
    off, tgt = grid_graph(rows, cols)
    for s in (0, 5 * cols + 7, rows * cols - 1):
        dist, reach, depth = oracle.bfs(off, tgt, s)
        r, c = divmod(np.arange(rows * cols), cols)
        sr, sc = divmod(s, cols)
        assert np.array_equal(dist, np.abs(r - sr) + np.abs(c - sc))


def test_bfs_petersen_and_networkx_fixtures():
    golden = {g["name"]: g for g in load("catalog.json")}
    for name, off, tgt in catalog():
        for s in range(len(off) - 1):
            dist, reach, depth = oracle.bfs(off, tgt, s)
            assert dist.tolist() == golden[name]["bfs"][s], (name, s)
            assert reach == sum(d >= 0 for d in dist)
            assert depth == max(dist)
    p = golden["petersen"]["bfs"][0]
    assert sorted(p) == [0, 1, 1, 1, 2, 2, 2, 2, 2, 2]


# ---- PageRank ----------------------------------------------------------------------


def test_pagerank_regular_and_edgeless_graphs_are_uniform():
    for n, edges in ((4, complete_edges(4)), (7, complete_edges(7)), (10, [(a, b) for a, b in
                     [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 5), (1, 6), (2, 7), (3, 8), (4, 9),
                      (5, 7), (7, 9), (9, 6), (6, 8), (8, 5)]]), (3, [])):
        off, tgt = undirected(n, edges)
        r, it, conv = oracle.pagerank(off, tgt, 0.85, 100, 1e-12)
        assert conv and np.allclose(r, 1.0 / n, atol=1e-15, rtol=0)


def test_pagerank_star_closed_form():
    """Star with k leaves: hub h and leaf l satisfy h = (1-d)/n + d*k*l,
    l = (1-d)/n + d*h/k (no dangling nodes)."""
    k, d = 8, 0.85
    n = k + 1
    a = np.array([[1.0, -d * k], [-d / k, 1.0]])
    h, leaf = np.linalg.solve(a, np.full(2, (1 - d) / n))
    off, tgt = undirected(n, [(0, i) for i in range(1, n)])
    r, _, conv = oracle.pagerank(off, tgt, d, 10000, 1e-15)
    assert conv
    assert abs(r[0] - h) < 1e-12 and np.allclose(r[1:], leaf, atol=1e-12, rtol=0)


def test_pagerank_fixed_point_matches_scipy_and_networkx():
    golden = {g["name"]: g for g in load("catalog.json")}
    for name, off, tgt in catalog():
        if len(off) == 1:
            continue
        r, _, conv = oracle.pagerank(off, tgt, 0.85, 10000, 1e-15)
        assert conv
        assert np.allclose(r, golden[name]["pagerank_scipy"], atol=1e-12, rtol=0), name
        assert np.allclose(r, golden[name]["pagerank_networkx"], atol=1e-9, rtol=0), name
        assert abs(r.sum() - 1.0) < 1e-12


def test_pagerank_iteration_semantics():
    """The count includes the converging iteration; max_iterations caps it; n == 0 is
    ({}, 0, converged) (pagerank.rs:56-63,115-183)."""
    off, tgt = undirected(4, complete_edges(4))
    r, it, conv = oracle.pagerank(off, tgt, 0.85, 100, 1e-6)
    assert (it, conv) == (1, True)  # uniform start is already the fixed point
    off, tgt = undirected(9, [(0, i) for i in range(1, 9)])
    _, it, conv = oracle.pagerank(off, tgt, 0.85, 3, 1e-12)
    assert (it, conv) == (3, False)
    r, it, conv = oracle.pagerank([0], [], 0.85, 10, 1e-6)
    assert (len(r), it, conv) == (0, 0, True)


# ---- k-truss -----------------------------------------------------------------------


def test_k_truss_complete_graphs():
    for n in range(3, 9):
        off, tgt = undirected(n, complete_edges(n))
        for k in range(3, n + 3):
            edges, it = oracle.k_truss(off, tgt, k)
            assert len(edges) == (math.comb(n, 2) if k <= n else 0), (n, k)


def test_k_truss_catalog_by_hand_and_networkx():
    by_hand = {
        ("wheel", 3): 12, ("wheel", 4): 0, ("two K4 joined by a bridge", 3): 12,
        ("two K4 joined by a bridge", 4): 12, ("diamond", 3): 5, ("diamond", 4): 0,
        ("K5", 5): 10, ("K5", 6): 0,
    }
    golden = {g["name"]: g for g in load("catalog.json")}
    for name, off, tgt in catalog():
        for k in range(3, 9):
            edges, _ = oracle.k_truss(off, tgt, k)
            assert [list(e) for e in edges] == sorted(golden[name]["k_truss"][str(k)]), (name, k)
            if (name, k) in by_hand:
                assert len(edges) == by_hand[(name, k)]


def test_k_truss_brute_force_and_fast_variant():
    rng = np.random.default_rng(8)
    for _ in range(150):
        n, edges = random_graph(rng, max_nodes=14, max_edges=60)
        off, tgt = undirected(n, edges)
        for k in (3, 4, 5):
            ref, it_ref = oracle.k_truss(off, tgt, k)
            assert ref == brute_truss(off, tgt, k)
            fast, it_fast = oracle.k_truss(off, tgt, k, fast=True)
            assert (fast, it_fast) == (ref, it_ref)


def test_k_truss_round_count_semantics():
    """iterations includes the final round that removes nothing; an edgeless graph gives 0."""
    off, tgt = undirected(5, complete_edges(5))
    assert oracle.k_truss(off, tgt, 5) == ([(a, b) for a, b in complete_edges(5)], 1)
    assert oracle.k_truss(off, tgt, 6)[1] == 2
    assert oracle.k_truss([0, 0, 0], [], 3) == ([], 0)
    # a self-loop is removed in round 1 even when it inflates round-1 supports
    off, tgt = undirected(3, [(0, 1), (1, 2), (0, 2), (0, 0)])
    assert oracle.k_truss(off, tgt, 3) == ([(0, 1), (0, 2), (1, 2)], 2)


def test_k_truss_fast_matches_ref_at_scale():
    off, tgt = er_graph(4096, 16, seed=1)
    er = load("er4096.json")
    import hashlib

    for k in (3, 4):
        ref, it = oracle.k_truss(off, tgt, k)
        assert oracle.k_truss(off, tgt, k, fast=True) == (ref, it)
        assert len(ref) == er["k_truss"][str(k)]["edges"]
        assert hashlib.sha256(np.asarray(ref, dtype="<u4").tobytes()).hexdigest() == er["k_truss"][str(k)]["sha256"]
    rm = load("rmat_counts.json")["rmat12_ef16_seed1"]
    off, tgt = rmat_graph(12, 16, seed=1)
    ref, it = oracle.k_truss(off, tgt, 4)
    assert oracle.k_truss(off, tgt, 4, fast=True) == (ref, it)
    assert len(ref) == rm["k_truss_4"]["edges"]
    assert hashlib.sha256(np.asarray(ref, dtype="<u4").tobytes()).hexdigest() == rm["k_truss_4"]["sha256"]


# ---- truss numbers ----------------------------------------------------------------


def by_truss(edges, truss, k):
    return [e for e, t in zip(edges, truss) if t >= k]


def test_truss_numbers_complete_graphs():
    for n in range(2, 9):
        off, tgt = undirected(n, complete_edges(n))
        edges, truss = oracle.truss_numbers(off, tgt)
        assert edges == [tuple(e) for e in complete_edges(n)]
        assert list(truss) == [max(n, 2)] * math.comb(n, 2), n
    edges, truss = oracle.truss_numbers([0, 0, 0], [])
    assert edges == [] and truss.size == 0


def test_truss_numbers_agree_with_k_truss_on_catalog_and_random_graphs():
    """truss(e) >= k exactly for the edges of the k-truss (oracle_k_truss_ref, itself pinned by
    networkx and brute-force peeling above), for every k up to one past the largest value."""
    cases = [(name, off, tgt) for name, off, tgt in catalog()]
    rng = np.random.default_rng(21)
    for i in range(120):
        n, edges = random_graph(rng, max_nodes=16, max_edges=70)
        off, tgt = undirected(n, edges)
        cases.append((f"random #{i}", off, tgt))
    for name, off, tgt in cases:
        edges, truss = oracle.truss_numbers(off, tgt)
        assert all(t >= 2 for t in truss), name
        top = int(max(truss)) if len(truss) else 2
        for k in range(3, top + 2):
            ref, _ = oracle.k_truss(off, tgt, k)
            assert by_truss(edges, truss, k) == ref, (name, k)
        if len(off) <= 12:
            for k in range(3, top + 2):
                assert by_truss(edges, truss, k) == brute_truss(off, tgt, k), (name, k)


def test_truss_numbers_by_hand():
    # wheel: every edge in the 3-truss, none in the 4-truss; two K4 joined by a bridge: the
    # bridge is in no triangle (2), the K4 edges are 4
    cat = {name: (off, tgt) for name, off, tgt in catalog()}
    edges, truss = oracle.truss_numbers(*cat["wheel"])
    assert set(int(t) for t in truss) == {3}
    edges, truss = oracle.truss_numbers(*cat["two K4 joined by a bridge"])
    assert sorted(int(t) for t in truss) == [2] + [4] * 12


def test_truss_numbers_rmat12_against_fast_k_truss():
    off, tgt = rmat_graph(12, 16, seed=1)
    edges, truss = oracle.truss_numbers(off, tgt)
    for k in (3, 4, 6, 9, int(truss.max()), int(truss.max()) + 1):
        ref, _ = oracle.k_truss(off, tgt, k, fast=True)
        assert by_truss(edges, truss, k) == ref, k


def test_truss_numbers_reject_directed_csr():
    with pytest.raises(ValueError):
        oracle.truss_numbers([0, 1, 1], [1])


def test_pagerank_first_iterates_by_hand_with_a_dangling_vertex():
    """A star 0-{1,2,3} plus an isolated vertex 4 (degree 0, dangling), d = 0.85, worked by hand
    from r0 = 1/5: D = r0[4] = 0.2, so every vertex gets (1-d)/5 + d*D/5 = 0.064, plus
    d * 3 * 0.2 / 1 = 0.51 at the centre and d * 0.2 / 3 at each leaf. The second iterate takes
    D = r1[4] = 0.064, the dangling mass of the input iterate (taking it from the output iterate
    would give another second iterate). L1 changes 0.748 and 0.49708 decide convergence."""
    off, tgt = undirected(5, [(0, 1), (0, 2), (0, 3)])
    r1, it, conv = oracle.pagerank(off, tgt, 0.85, 1, 0.0)
    assert it == 1 and not conv
    assert np.allclose(r1, [0.574, 0.03 + 6.8 / 75, 0.03 + 6.8 / 75, 0.03 + 6.8 / 75, 0.064], rtol=0, atol=1e-15)
    r2, it, conv = oracle.pagerank(off, tgt, 0.85, 2, 0.0)
    assert it == 2 and not conv
    leaf2 = 0.03 + 0.85 * (0.574 / 3 + 0.0128)
    assert np.allclose(r2, [0.34858, leaf2, leaf2, leaf2, 0.04088], rtol=0, atol=1e-14)
    assert abs(leaf2 - 0.20351333333333333) < 1e-15
    # L1 change of the first iterate 0.748, of the second 0.49708
    r, it, conv = oracle.pagerank(off, tgt, 0.85, 10, 0.75)
    assert (it, conv) == (1, True) and np.allclose(r, r1, rtol=0, atol=1e-15)
    r, it, conv = oracle.pagerank(off, tgt, 0.85, 10, 0.5)
    assert (it, conv) == (2, True) and np.allclose(r, r2, rtol=0, atol=1e-14)
    # third change 0.4146572: a tolerance between the second and third changes stops at 3
    r, it, conv = oracle.pagerank(off, tgt, 0.85, 10, 0.497)
    assert (it, conv) == (3, True)


def test_triangle_sample_phases_add_up_to_the_full_count():
    """The CPU baseline's sampled count: the phases of one stride partition the DAG rows."""
    for name, off, tgt in [("ER-4096", *er_graph(4096, 16, 1)), ("RMAT-12", *rmat_graph(12, 16, 1))]:
        full = oracle.triangle_count(off, tgt, fast=True)
        sample = oracle.TriangleSample(off, tgt)
        try:
            assert sample.count_rows(1) == full, name
            for stride in (2, 7, 64):
                assert sum(sample.count_rows(stride, p) for p in range(stride)) == full, (name, stride)
        finally:
            sample.close()
    assert full == oracle.triangle_count(off, tgt)
