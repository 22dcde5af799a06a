"""GPU results against the CPU oracle on the same CSR.

Integer results (triangle counts, BFS distances, k-truss edge sets and round counts) must be
bit-exact; PageRank (FP32 ranks, FP64 reductions) must be within L1 <= 1e-6 of the FP64 oracle
run for the same number of iterations (the tolerance of the graph-kernel path's target).
Comparisons follow crates/grafeo-cuda-algos/tests/correctness.rs:234-326 (BFS from every source,
k = 3..8 for k-truss, triangle counts through both the symmetric and the oriented builder).
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import grafeo_cuda
import oracle
from grafeo_cuda import _native
from graphs import catalog, er_graph, grid_graph, random_graph, rmat_graph, undirected

pytestmark = pytest.mark.gpu

PR_L1_TOL = 1e-6


def upload(device, off, tgt):
    return grafeo_cuda.GpuGraph.from_csr(off, tgt, device)


def check_triangles(name, device, off, tgt, g):
    expected = oracle.triangle_count(off, tgt)
    assert g.triangle_count()["total_triangles"] == expected, f"{name}: triangle_count"
    # the bitmap kernel's masks of the top ids: off, and at 128 ids (masks on short lists)
    for top in ("0", "128"):
        os.environ["GC_TC_TOP"] = top
        try:
            assert g.triangle_count()["total_triangles"] == expected, f"{name}: triangle_count (top={top})"
        finally:
            del os.environ["GC_TC_TOP"]
    # every triangle method (hub bitmap, hash probes, binary-search lists, middle-vertex
    # wedges) must agree
    for method in ("hash", "lists", "wedge"):
        os.environ["GC_TC_METHOD"] = method
        try:
            assert g.triangle_count()["total_triangles"] == expected, f"{name}: triangle_count ({method})"
        finally:
            del os.environ["GC_TC_METHOD"]
    n = len(off) - 1
    pairs = [(u, int(v)) for u in range(n) for v in tgt[off[u]:off[u + 1]] if v > u]
    o_off, o_tgt = _native.from_edge_pairs_oriented(pairs, n)
    og = upload(device, o_off, o_tgt)
    assert og.triangle_count_oriented()["total_triangles"] == expected, f"{name}: oriented"
    assert og.triangle_count_oriented()["total_triangles"] == oracle.triangle_count_oriented(o_off, o_tgt)


def check_tc_partitions(name, g, expected, monkeypatch):
    """The bitmap kernel splits the middle vertices into parts (one per XCD or pair of XCDs) and
    runs short rows whole, masks the top ids of long lists and walks short lists a quarter wave
    each: every split, mask width and walk must give the same count."""
    for top in ("0", "128", "1024", "8192"):
        for parts in ("1", "2", "3", "8"):
            for split in ("0", "64", "4294967295"):
                monkeypatch.setenv("GC_TC_TOP", top)
                monkeypatch.setenv("GC_TC_PARTS", parts)
                monkeypatch.setenv("GC_TC_SPLIT", split)
                assert g.triangle_count()["total_triangles"] == expected, \
                    f"{name}: top={top} parts={parts} split={split}"
    monkeypatch.delenv("GC_TC_TOP")
    monkeypatch.delenv("GC_TC_PARTS")
    monkeypatch.delenv("GC_TC_SPLIT")
    # the short hot parts a quarter wave each, one to four octs per lane and step, up to a few
    # list lengths (0: a wave step per list)
    for quarters, qmax in (("0", None), ("1", "1"), ("1", "63"), ("2", "8"), ("3", "48"), ("4", "4096")):
        monkeypatch.setenv("GC_TC_QUARTERS", quarters)
        if qmax is None:
            monkeypatch.delenv("GC_TC_QMAX", raising=False)
        else:
            monkeypatch.setenv("GC_TC_QMAX", qmax)
        assert g.triangle_count()["total_triangles"] == expected, f"{name}: quarters={quarters} qmax={qmax}"
    monkeypatch.delenv("GC_TC_QUARTERS")
    monkeypatch.delenv("GC_TC_QMAX")
    # four-wave workgroups (each wave's bitmap at its own LDS base) against the one-wave default
    monkeypatch.setenv("GC_TC_WG", "4")
    assert g.triangle_count()["total_triangles"] == expected, f"{name}: four-wave workgroups"
    monkeypatch.delenv("GC_TC_WG")
    # the middle-vertex side (tc_mid16) off; taking every edge it may (per-edge cost 0); only the
    # edges whose walk is empty (a cost no walk beats); in units of 64 entries; with the masks,
    # parts and quarter walks it shares with tc_bitmap16 changed
    for mid, cost, unit, top, parts, quarters in (("0", None, None, None, None, None),
                                                  ("1", "0", None, None, None, None),
                                                  ("1", "4294967295", None, None, None, None),
                                                  ("1", "0", "64", None, None, None),
                                                  ("1", "1", "64", "0", "3", None),
                                                  ("1", "0", None, "1024", "8", "0"),
                                                  ("1", "2", "100", "128", "1", "1")):
        for var, val in (("GC_TC_MID", mid), ("GC_TC_MID_COST", cost), ("GC_TC_MID_UNIT", unit),
                         ("GC_TC_TOP", top), ("GC_TC_PARTS", parts), ("GC_TC_QUARTERS", quarters)):
            if val is None:
                monkeypatch.delenv(var, raising=False)
            else:
                monkeypatch.setenv(var, val)
        assert g.triangle_count()["total_triangles"] == expected, \
            f"{name}: mid={mid} cost={cost} unit={unit} top={top} parts={parts} quarters={quarters}"
    for var in ("GC_TC_MID", "GC_TC_MID_COST", "GC_TC_MID_UNIT", "GC_TC_TOP", "GC_TC_PARTS", "GC_TC_QUARTERS"):
        monkeypatch.delenv(var, raising=False)


def check_pagerank_kernels(name, off, tgt, g, monkeypatch):
    """Every SpMV kernel against the oracle: the default (push-pull on a symmetric graph with the
    binned cold entries and the hot stream sliced one row per lane, with its push and hot sets at
    their defaults, at a few dozen ids, and at one id; with pieces of one or two groups of eight
    entries, so that the wide slices are cut and finished from their pieces; and with small
    staging buffers and bins, so that the scatter takes many sub-ranges and the rows many bins),
    the same bins with the hot stream in 512-entry chunks (bin), the nonzero-split pull that
    gathers the cold entries (merge), the row-tiled push-pull (push), the pull kernel with its hot set in LDS (default size
    and a set smaller than the graph), and one workgroup per row tile."""
    pr_vars = ("GC_PR_KERNEL", "GC_PR_PUSH", "GC_PR_HOT", "GC_PR_STAGE", "GC_PR_BIN_ROWS", "GC_PR_BIN_VALUES",
               "GC_PR_SLICE_CAP")
    cases = ((None, None, None), (None, "64", "96"), (None, "1", "1"), (None, "37", None),
             ("bin", "1", "1", "512", "1", "8"), ("bin", "3", "40", "512", "7", "100"),
             ("bin", None, None, "1000", "64", "64"),
             ("slice", None, None), ("slice", "64", "96", None, None, None, "1"), ("slice", "1", "1", "512", "1", "8"),
             ("slice", "3", "40", "512", "7", "100", "2"),
             ("merge", None, None), ("merge", "64", "96"), ("merge", "1", "1"), ("merge", "37", None),
             ("push", None, None), ("push", "37", None),
             ("rows", None, None), ("hot", None, "1024"), ("hot", None, None))
    for case in cases:
        values = case + (None,) * (len(pr_vars) - len(case))
        for var, value in zip(pr_vars, values):
            if value is None:
                monkeypatch.delenv(var, raising=False)
            else:
                monkeypatch.setenv(var, value)
        label = ", ".join(f"{v}={x}" for v, x in zip(pr_vars, values) if x is not None)
        check_pagerank(f"{name} ({label or 'defaults'})", off, tgt, g)
    for var in pr_vars:
        monkeypatch.delenv(var, raising=False)


def check_bfs(name, off, tgt, g, source):
    dist, reach, depth = oracle.bfs(off, tgt, source)
    r = g.bfs_numpy(source)
    assert np.array_equal(r["distances"], dist), f"{name}: BFS from {source}"
    assert r["reachable"] == reach, f"{name}: reachable from {source}"
    assert r["max_depth"] == depth, f"{name}: max depth from {source}"


def check_pagerank(name, off, tgt, g):
    r = g.pagerank_numpy(0.85, 1000, 1e-7)
    n = len(off) - 1
    assert len(r["ranks"]) == n
    if n == 0:
        return
    ref, it, _ = oracle.pagerank(off, tgt, 0.85, r["iterations"], 0.0)
    assert it == r["iterations"]
    l1 = float(np.abs(r["ranks"].astype(np.float64) - ref).sum())
    assert l1 <= PR_L1_TOL, f"{name}: L1 {l1}"
    fixed, _, conv = oracle.pagerank(off, tgt, 0.85, 1000, 1e-12)
    assert np.max(np.abs(r["ranks"] - fixed)) < 1e-4, f"{name}: far from the fixed point"
    assert abs(float(r["ranks"].astype(np.float64).sum()) - 1.0) < 1e-3


def check_k_truss(name, off, tgt, g, monkeypatch, ks=range(3, 9), both_paths=True, fast_oracle=False):
    """Both GPU paths (the DAG rounds, with the supports from the bitmap walk with and without its
    middle side (every edge it may take at cost 0, units of 64 entries) and from the wedge
    enumeration, and the per-entry rounds) against the oracle."""
    paths = ((("0", None, None, None, None), ("0", None, "0", None, None), ("0", None, None, "0", "64"),
              ("0", "wedge", None, None, None), ("1", None, None, None, None)) if both_paths
             else (("0", None, None, None, None), ("0", None, "0", None, None)))
    names = ("GC_KTRUSS_DIRECT", "GC_KTRUSS_SUPPORT", "GC_KTRUSS_MID", "GC_KTRUSS_MID_COST", "GC_KTRUSS_MID_UNIT")
    for k in ks:
        edges, iterations = oracle.k_truss(off, tgt, k, fast=fast_oracle)
        for path in paths:
            for var, val in zip(names, path):
                if val is None:
                    monkeypatch.delenv(var, raising=False)
                else:
                    monkeypatch.setenv(var, val)
            r = g.k_truss(k)
            assert r["edges"] == edges, f"{name}: {k}-truss ({dict(zip(names, path))})"
            assert r["iterations"] == iterations, f"{name}: {k}-truss rounds ({dict(zip(names, path))})"
            assert r["k"] == k
    for var in names:
        monkeypatch.delenv(var, raising=False)


def check_msbfs(name, off, tgt, g, sources):
    r = g.msbfs_numpy(sources)
    assert r["distances"].shape == (len(sources), len(off) - 1)
    for row, source in enumerate(sources):
        dist, reach, depth = oracle.bfs(off, tgt, int(source))
        assert np.array_equal(r["distances"][row], dist), f"{name}: MS-BFS row {row} (source {source})"
        assert r["reachable"][row] == reach and r["max_depth"][row] == depth, f"{name}: MS-BFS {source}"


def check_truss_numbers(name, off, tgt, g, monkeypatch=None):
    """Trussness against the oracle's peeling; with monkeypatch, also from the wedge supports."""
    edges, truss = oracle.truss_numbers(off, tgt)
    for support in ((None, "wedge", "nomid") if monkeypatch is not None else (None,)):
        if support == "nomid":  # the bitmap walk's supports without its middle side
            monkeypatch.delenv("GC_KTRUSS_SUPPORT", raising=False)
            monkeypatch.setenv("GC_KTRUSS_MID", "0")
        elif support is not None:
            monkeypatch.setenv("GC_KTRUSS_SUPPORT", support)
        r = g.truss_numbers_numpy()
        assert [(int(u), int(v)) for u, v in r["edges"]] == edges, f"{name}: truss edges"
        assert np.array_equal(r["truss"], truss), f"{name}: truss numbers (support={support})"
        assert r["max_truss"] == (int(truss.max()) if truss.size else 0), f"{name}: max truss"
    if monkeypatch is not None:
        monkeypatch.delenv("GC_KTRUSS_SUPPORT", raising=False)
        monkeypatch.delenv("GC_KTRUSS_MID", raising=False)


def test_catalog(device, monkeypatch):
    for name, off, tgt in catalog():
        g = upload(device, off, tgt)
        check_triangles(name, device, off, tgt, g)
        for source in range(len(off) - 1):
            check_bfs(name, off, tgt, g, source)
        check_pagerank(name, off, tgt, g)
        check_k_truss(name, off, tgt, g, monkeypatch)
        check_truss_numbers(name, off, tgt, g, monkeypatch)
        check_msbfs(name, off, tgt, g, list(range(len(off) - 1)))


def test_random_graphs(device, monkeypatch):
    """The reference's proptest (48 cases, n < 16, <= 48 pairs with loops), run 96 times."""
    rng = np.random.default_rng(20240611)
    cases = [(1, [])] + [random_graph(rng) for _ in range(95)]
    for case, (n, edges) in enumerate(cases):
        off, tgt = undirected(n, edges)
        g = upload(device, off, tgt)
        name = f"random #{case} (n={n}, edges={edges})"
        check_triangles(name, device, off, tgt, g)
        check_bfs(name, off, tgt, g, 0)
        check_bfs(name, off, tgt, g, n - 1)
        check_pagerank(name, off, tgt, g)
        with monkeypatch.context() as mp:
            # one pushed hub, so the nonzero-split pull covers the other rows in one short chunk
            mp.setenv("GC_PR_KERNEL", "merge")
            mp.setenv("GC_PR_PUSH", "1")
            mp.setenv("GC_PR_HOT", "1")
            check_pagerank(name + " (merge, push=1)", off, tgt, g)
            # the same with the cold entries binned, one row per bin
            mp.setenv("GC_PR_KERNEL", "bin")
            mp.setenv("GC_PR_BIN_ROWS", "1")
            check_pagerank(name + " (bin, push=1, one row per bin)", off, tgt, g)
            # the hot stream sliced, slices cut into pieces of one group
            mp.setenv("GC_PR_KERNEL", "slice")
            mp.setenv("GC_PR_SLICE_CAP", "1")
            check_pagerank(name + " (slice, push=1, one row per bin, pieces of one group)", off, tgt, g)
        check_k_truss(name, off, tgt, g, monkeypatch, ks=(3, 4, 5))
        check_truss_numbers(name, off, tgt, g, monkeypatch)
        check_msbfs(name, off, tgt, g, list(range(n)) + [0, n - 1])


def test_directed_csr_semantics(device, monkeypatch):
    """Non-symmetric CSRs take the reference's per-entry definitions (no relabelled DAG)."""
    rng = np.random.default_rng(7)
    for case in range(40):
        n, edges = random_graph(rng, max_nodes=14, max_edges=40)
        off, tgt = _native.from_edge_pairs(edges, n, False)
        g = upload(device, off, tgt)
        name = f"directed #{case}"
        assert g.triangle_count()["total_triangles"] == oracle.triangle_count(off, tgt), name
        assert g.triangle_count_oriented()["total_triangles"] == oracle.triangle_count_oriented(off, tgt), name
        for source in range(n):
            check_bfs(name, off, tgt, g, source)
        check_msbfs(name, off, tgt, g, list(range(n)))
        r = g.pagerank_numpy(0.85, 30, 0.0)
        ref, _, _ = oracle.pagerank(off, tgt, 0.85, 30, 0.0)
        assert float(np.abs(r["ranks"] - ref).sum()) <= PR_L1_TOL, name
        for k in (3, 4):
            edges_o, it_o = oracle.k_truss(off, tgt, k)
            r = g.k_truss(k)
            assert (r["edges"], r["iterations"]) == (edges_o, it_o), f"{name} k={k}"
        if not g.is_symmetric() and len(tgt):
            with pytest.raises(RuntimeError, match="symmetric"):
                g.truss_numbers()


def test_er_4096(device, monkeypatch):
    """Triangle count, plain method: Erdos-Renyi n=4096, average degree 16."""
    off, tgt = er_graph(4096, 16, seed=1)
    g = upload(device, off, tgt)
    check_triangles("ER-4096", device, off, tgt, g)
    check_tc_partitions("ER-4096", g, oracle.triangle_count(off, tgt), monkeypatch)
    check_pagerank_kernels("ER-4096", off, tgt, g, monkeypatch)
    for source in (0, 17, 4095):
        check_bfs("ER-4096", off, tgt, g, source)
    check_pagerank("ER-4096", off, tgt, g)
    check_k_truss("ER-4096", off, tgt, g, monkeypatch, ks=(3, 4))
    check_truss_numbers("ER-4096", off, tgt, g, monkeypatch)
    # 130 sources: one full traversal of 128 and one of 2, with repeats
    sources = [int(x) for x in np.random.default_rng(4096).integers(0, 4096, size=127)] + [0, 0, 4095]
    check_msbfs("ER-4096", off, tgt, g, sources)


@pytest.mark.parametrize("scale", [10, 14])
def test_rmat_small(device, monkeypatch, scale):
    off, tgt = rmat_graph(scale, 16, seed=1)
    g = upload(device, off, tgt)
    check_triangles(f"RMAT-{scale}", device, off, tgt, g)
    check_tc_partitions(f"RMAT-{scale}", g, oracle.triangle_count(off, tgt, fast=True), monkeypatch)
    check_pagerank_kernels(f"RMAT-{scale}", off, tgt, g, monkeypatch)
    n = len(off) - 1
    for source in np.random.default_rng(scale).integers(0, n, size=4):
        check_bfs(f"RMAT-{scale}", off, tgt, g, int(source))
    check_pagerank(f"RMAT-{scale}", off, tgt, g)
    # the per-entry path merges full lists (hub x hub), so it is checked at the small scale only
    check_k_truss(f"RMAT-{scale}", off, tgt, g, monkeypatch, ks=(3, 4, 6), both_paths=scale <= 10,
                  fast_oracle=scale > 10)
    check_truss_numbers(f"RMAT-{scale}", off, tgt, g, monkeypatch)
    check_msbfs(f"RMAT-{scale}", off, tgt, g, [int(x) for x in np.random.default_rng(scale).integers(0, n, size=128)])


@pytest.mark.parametrize("scale", [16, 18])
def test_rmat_tc_sides(device, monkeypatch, scale):
    """The triangle count's three sides on graphs with ids below the top 32768 (R-MAT 16: the
    16-bit middle side over the top half, the 32-bit tier over the rest) and below the top 131072
    (R-MAT 18: rows on no middle side too): each side off, every edge it may take (per-edge cost
    0), only the empty walks, small work units, and the parts and masks changed."""
    off, tgt = rmat_graph(scale, 16, seed=5)
    expected = oracle.triangle_count(off, tgt, fast=True)
    g = upload(device, off, tgt)
    assert g.triangle_count()["total_triangles"] == expected, f"RMAT-{scale}: default"
    names = ("GC_TC_MID", "GC_TC_MID32", "GC_TC_MID_COST", "GC_TC_MID_UNIT", "GC_TC_PARTS", "GC_TC_TOP")
    for case in (("0", None, None, None, None, None), ("1", "0", None, None, None, None),
                 ("1", "1", None, None, None, None),
                 ("1", "1", "0", None, None, None), ("1", "1", "4294967295", None, None, None),
                 ("1", "1", "0", "64", "3", None), ("1", "0", "1", "100", "8", "0"),
                 ("1", "1", "2", "64", "1", "1024")):
        for var, val in zip(names, case):
            if val is None:
                monkeypatch.delenv(var, raising=False)
            else:
                monkeypatch.setenv(var, val)
        assert g.triangle_count()["total_triangles"] == expected, f"RMAT-{scale}: {dict(zip(names, case))}"
    for var in names:
        monkeypatch.delenv(var, raising=False)
    # the 32-bit tier over every row below the hot ids, and over a band narrower than the default
    monkeypatch.setenv("GC_TC_MID32", "1")
    for ids in ("4294967295", "40000"):
        for cost in ("0", "2"):
            monkeypatch.setenv("GC_TC_MID32_IDS", ids)
            monkeypatch.setenv("GC_TC_MID_COST", cost)
            assert g.triangle_count()["total_triangles"] == expected, f"RMAT-{scale}: ids={ids} cost={cost}"
    monkeypatch.delenv("GC_TC_MID32_IDS")
    monkeypatch.delenv("GC_TC_MID_COST")
    # the middle kernels' units taken one and three per atomic
    for grab in ("1", "3"):
        monkeypatch.setenv("GC_TC_MID_GRAB", grab)
        monkeypatch.setenv("GC_TC_MID_UNIT", "64")
        assert g.triangle_count()["total_triangles"] == expected, f"RMAT-{scale}: grab={grab}"
    for var in ("GC_TC_MID_GRAB", "GC_TC_MID_UNIT", "GC_TC_MID32"):
        monkeypatch.delenv(var)
    # the middle kernel's short walks 4, 8 or 16 lanes each or a quarter wave (0), up to 4096 octs;
    # the cold parts' walk in tc_bitmap16 4, 8 or 16 lanes per part
    for group, cold, qmax in (("0", "4", None), ("4", "8", None), ("8", "16", "4096"), ("16", "4", "1")):
        monkeypatch.setenv("GC_TC_MID_GROUP", group)
        monkeypatch.setenv("GC_TC_COLD_GROUP", cold)
        if qmax is None:
            monkeypatch.delenv("GC_TC_QMAX", raising=False)
        else:
            monkeypatch.setenv("GC_TC_QMAX", qmax)
        assert g.triangle_count()["total_triangles"] == expected, \
            f"RMAT-{scale}: mid group {group} cold group {cold} qmax {qmax}"
    for var in ("GC_TC_MID_GROUP", "GC_TC_COLD_GROUP", "GC_TC_QMAX"):
        monkeypatch.delenv(var, raising=False)
    # the k-truss supports with ids below the hot ones: the bitmap walk with and without its middle side
    kedges, kit = oracle.k_truss(off, tgt, 4, fast=True)
    for mid, cost in (("1", None), ("0", None), ("1", "0")):
        monkeypatch.setenv("GC_KTRUSS_MID", mid)
        if cost is None:
            monkeypatch.delenv("GC_KTRUSS_MID_COST", raising=False)
        else:
            monkeypatch.setenv("GC_KTRUSS_MID_COST", cost)
        r = g.k_truss(4)
        assert r["edges"] == kedges and r["iterations"] == kit, f"RMAT-{scale}: 4-truss mid={mid} cost={cost}"
    monkeypatch.delenv("GC_KTRUSS_MID")
    monkeypatch.delenv("GC_KTRUSS_MID_COST", raising=False)
    # every row taken singly, and none (rows in chunks only)
    for heavy in ("1", "4294967295"):
        monkeypatch.setenv("GC_TC_HEAVY_DEGREE", heavy)
        assert g.triangle_count()["total_triangles"] == expected, f"RMAT-{scale}: heavy degree {heavy}"
    monkeypatch.delenv("GC_TC_HEAVY_DEGREE")


def test_bfs_schedule_switches(device, monkeypatch):
    """Distances under other fixed-rule settings against the oracle: bottom-up levels launched
    ahead or one per read-back (GC_BFS_AHEAD), Beamer's top-down -> bottom-up threshold at 1
    (bottom-up as soon as allowed), 14 (the default) and 1000 (almost never), and the bottom-up
    tail after Beamer's rule would return at 0 and 3 levels (GC_BFS_TAIL; with 0 one traversal
    can start several bottom-up runs, each clearing the runs' counter sets first), on the catalog,
    a grid (many levels) and R-MAT 14 (hubs). The schedule changes which kernels run, never
    the distances."""
    graphs = [(name, off, tgt) for name, off, tgt in catalog()]
    graphs.append(("grid 40x30",) + tuple(grid_graph(40, 30)))
    graphs.append(("RMAT-14",) + tuple(rmat_graph(14, 16, seed=3)))
    for name, off, tgt in graphs:
        g = upload(device, off, tgt)
        n = len(off) - 1
        sources = sorted({0, n // 2, n - 1}) if n else []
        expected = {s: oracle.bfs(off, tgt, s) for s in sources}
        for ahead in ("0", "1"):
            for alpha in ("1", "14", "1000"):
                for tail in ("0", "3"):
                    monkeypatch.setenv("GC_BFS_AHEAD", ahead)
                    monkeypatch.setenv("GC_BFS_ALPHA", alpha)
                    monkeypatch.setenv("GC_BFS_TAIL", tail)
                    for s in sources:
                        dist, reach, depth = expected[s]
                        r = g.bfs_numpy(s)
                        tag = f"{name} from {s} (ahead={ahead}, alpha={alpha}, tail={tail})"
                        assert np.array_equal(r["distances"], dist), tag
                        assert (r["reachable"], r["max_depth"]) == (reach, depth), tag
    monkeypatch.delenv("GC_BFS_AHEAD")
    monkeypatch.delenv("GC_BFS_ALPHA")
    monkeypatch.delenv("GC_BFS_TAIL")


def test_msbfs_deep_graphs(device):
    """Multi-source BFS deeper than 254 levels: the levels kept as bytes are written out as
    distances at level 255 and the later levels written directly (a 600-vertex path, depth up to
    599, and a 300 x 2 grid, depth up to 300), against the oracle."""
    path = [(i, i + 1) for i in range(599)]
    for name, (off, tgt) in (("path 600", undirected(600, path)), ("grid 300x2", grid_graph(300, 2))):
        g = upload(device, off, tgt)
        n = len(off) - 1
        check_msbfs(name, off, tgt, g, [0, n // 3, n // 2, n - 1])


def test_pagerank_layout_sizes(device, monkeypatch):
    """The binned layout's reported sizes (gc_pagerank_layout) against counts taken here from
    the degree-ordered CSR: cold entries (entries of the pull rows below the hot set), pull rows,
    cold sources, the hub-piece entries (entries of the push-set rows inside the push set) and
    the 16-bit stream (each pull row's entries in the hot set, at least one)."""
    from ctypes import c_uint64

    from grafeo_cuda._native import check, lib

    off, tgt = rmat_graph(14, 16, seed=1)
    g = upload(device, off, tgt)
    for push, hot in ((None, None), ("64", "96"), ("1", "1")):
        for var, value in (("GC_PR_PUSH", push), ("GC_PR_HOT", hot)):
            if value is None:
                monkeypatch.delenv(var, raising=False)
            else:
                monkeypatch.setenv(var, value)
        check_pagerank(f"RMAT-14 layout (push={push}, hot={hot})", off, tgt, g)
        vals = (c_uint64 * 19)()
        check(lib.gc_pagerank_layout(g._handle, vals, 19))
        lay = dict(zip(("push", "hot", "stage", "bin_rows", "bin_values", "cold", "values", "slot_rows",
                        "stream", "chunks", "subs", "bins", "runs", "pieces", "piece_entries", "pull",
                        "sources", "build_ns", "grid"), (int(v) for v in vals)))
        rl = g.relabelled_numpy()
        roff, rtgt = rl["offsets"].astype(np.int64), rl["targets"].astype(np.int64)
        n = roff.size - 1
        pb, hb = n - lay["push"], n - lay["hot"]
        deg = np.diff(roff)
        z = min(int(np.count_nonzero(deg == 0)), pb)
        row = np.repeat(np.arange(n), deg)
        pull = (row >= z) & (row < pb)
        assert lay["cold"] == int(np.count_nonzero(pull & (rtgt < hb)))
        assert lay["pull"] == pb - z and lay["sources"] == max(hb - z, 0)
        assert lay["piece_entries"] == int(np.count_nonzero((row >= pb) & (rtgt >= pb)))
        hot_per_row = np.bincount(row[pull & (rtgt >= hb)], minlength=n)[z:pb]
        assert lay["stream"] == int(np.maximum(hot_per_row, 1).sum())
        assert lay["chunks"] == (lay["stream"] + 511) // 512
        assert lay["values"] >= lay["cold"] and lay["values"] % 8 == 0 and lay["slot_rows"] % 8 == 0
        assert lay["grid"] >= 1
        # the sliced hot stream: 64 rows per slice, each as wide as its widest row's hot entries,
        # in groups of eight and pairs
        monkeypatch.setenv("GC_PR_KERNEL", "slice")
        monkeypatch.setenv("GC_PR_SLICE_CAP", "4")
        check_pagerank(f"RMAT-14 sliced layout (push={push}, hot={hot})", off, tgt, g)
        monkeypatch.delenv("GC_PR_KERNEL")
        monkeypatch.delenv("GC_PR_SLICE_CAP")
        vals = (c_uint64 * 25)()
        check(lib.gc_pagerank_layout(g._handle, vals, 25))
        sliced, slices, entries, items, wide, slots = (int(v) for v in vals[19:25])
        assert sliced == 1 and slices == (pb - z + 63) // 64
        width = np.zeros(slices * 64, np.int64)
        width[:pb - z] = hot_per_row
        width = width.reshape(slices, 64).max(axis=1)
        width += width & 1
        assert entries == int((512 * (width // 8) + 64 * (width % 8)).sum())
        cut = width // 8 > 4  # wider than pieces of 4 groups
        assert wide == int(np.count_nonzero(cut))
        assert slots == int(((width[cut] // 8 + 3) // 4).sum())
        assert items == slices - wide + slots + lay["pieces"]
