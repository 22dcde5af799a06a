"""Known-answer GPU tests ported from crates/grafeo-cuda-algos/tests/roundtrip.rs and the spot
checks of crates/grafeo-cuda-algos/tests/correctness.rs:330-386. They run through the C ABI via
the `grafeo_cuda` module on the device."""

from __future__ import annotations

import numpy as np
import pytest

import grafeo_cuda
from grafeo_cuda import _native
from graphs import complete_edges, path_edges, undirected

pytestmark = pytest.mark.gpu

K4 = ([0, 3, 6, 9, 12], [1, 2, 3, 0, 2, 3, 0, 1, 3, 0, 1, 2])
P5 = ([0, 1, 3, 5, 7, 8], [1, 0, 2, 1, 3, 2, 4, 3])


def graph(device, offsets, targets):
    return grafeo_cuda.GpuGraph.from_csr(offsets, targets, device)


def test_native_library_is_the_one_loaded():
    import os

    assert os.path.samefile(_native.LIB_PATH, os.path.join(os.path.dirname(grafeo_cuda.__file__), "libgrafeo_hip.so"))
    assert _native.device_count() >= 1


def test_device_properties(device):
    assert device.name
    assert device.total_memory > 0


def test_triangle_count_k4(device):
    assert graph(device, *K4).triangle_count()["total_triangles"] == 4


def test_triangle_count_path(device):
    assert graph(device, *P5).triangle_count()["total_triangles"] == 0


def test_bfs_k4(device):
    r = graph(device, *K4).bfs(0)
    assert r["reachable"] == 4
    assert r["max_depth"] == 1
    assert r["distances"] == [0, 1, 1, 1]


def test_bfs_path(device):
    r = graph(device, *P5).bfs(0)
    assert r["reachable"] == 5
    assert r["max_depth"] == 4
    assert r["distances"] == [0, 1, 2, 3, 4]


def test_pagerank_k4(device):
    r = graph(device, *K4).pagerank(0.85, 100, 1e-6)
    assert r["converged"]
    assert all(abs(x - 0.25) < 0.01 for x in r["ranks"])


def test_empty_graph(device):
    g = graph(device, [0], [])
    assert (g.num_nodes, g.num_edges) == (0, 0)
    assert g.triangle_count()["total_triangles"] == 0
    assert g.triangle_count_oriented()["total_triangles"] == 0
    pr = g.pagerank(0.85, 20, 1e-6)
    assert pr["ranks"] == [] and pr["converged"] and pr["iterations"] == 0
    assert g.k_truss(3)["edges"] == []
    with pytest.raises(RuntimeError, match="out of range") as info:
        g.bfs(0)
    assert info.value.kind == "Kernel"


def test_edgeless_graph(device):
    g = graph(device, [0, 0, 0, 0], [])
    assert g.triangle_count()["total_triangles"] == 0
    assert g.triangle_count_oriented()["total_triangles"] == 0
    kt = g.k_truss(3)
    assert kt["edges"] == [] and kt["iterations"] == 0
    r = g.bfs(1)
    assert r["distances"] == [-1, 0, -1]
    assert (r["reachable"], r["max_depth"]) == (1, 0)
    # every node is dangling: its rank is spread evenly instead of leaking
    pr = g.pagerank(0.85, 50, 1e-7)
    assert all(abs(x - 1.0 / 3.0) < 1e-5 for x in pr["ranks"])


def test_bfs_wide_frontier_after_first_level(device):
    """roundtrip.rs:186-222: node 1 fans out to 500,000 chains of 4, so levels 2..5 are each
    500,000 wide (wider than the GPU keeps resident)."""
    width, depth = 500_000, 4
    chain = np.arange(width, dtype=np.int64)
    edges = [np.array([[0, 1]])]
    edges.append(np.stack([np.ones(width, np.int64), 2 + chain], axis=1))
    for level in range(3, depth + 2):
        edges.append(np.stack([2 + (level - 3) * width + chain, 2 + (level - 2) * width + chain], axis=1))
    pairs = np.concatenate(edges)
    n = 2 + depth * width
    expected = np.concatenate([[0, 1]] + [np.full(width, level) for level in range(2, depth + 2)])
    off, tgt = _native.from_edge_pairs(pairs, n, True)
    g = graph(device, off, tgt)
    for run in range(5):
        r = g.bfs_numpy(0)
        wrong = int(np.count_nonzero(r["distances"] != expected))
        assert wrong == 0, f"run {run}: {wrong} nodes have a wrong distance"
        assert r["reachable"] == n
        assert r["max_depth"] == depth + 1


def test_triangle_count_oriented_k5(device):
    off, tgt = _native.from_edge_pairs_oriented(complete_edges(5), 5)
    assert graph(device, off, tgt).triangle_count_oriented()["total_triangles"] == 10


@pytest.mark.parametrize(
    "offsets,targets",
    [([], []), ([0, 5], [1]), ([0, 1, 2], [7, 0]), ([0, 2, 1, 2], [1, 2]), ([0, 2, 2], [1, 0])],
)
def test_invalid_csr_is_rejected_before_transfer(device, offsets, targets):
    with pytest.raises(RuntimeError, match="invalid CSR") as info:
        graph(device, offsets, targets)
    assert info.value.kind == "InvalidCsr"


def test_argument_errors(device):
    g = graph(device, *K4)
    with pytest.raises(RuntimeError, match="out of range"):
        g.bfs(4)
    with pytest.raises(RuntimeError, match="k must be >= 3"):
        g.k_truss(2)


def test_spot_checks(device):
    """correctness.rs:330-386: K4 = 4, K5 = 10 triangles; path of 6 from 0: depth 5,
    6 reachable; K5 5-truss has all 10 edges, 6-truss none."""
    assert graph(device, *undirected(4, complete_edges(4))).triangle_count()["total_triangles"] == 4
    k5 = graph(device, *undirected(5, complete_edges(5)))
    assert k5.triangle_count()["total_triangles"] == 10
    assert len(k5.k_truss(5)["edges"]) == 10
    assert k5.k_truss(6)["edges"] == []
    p = graph(device, *undirected(6, path_edges(6))).bfs(0)
    assert (p["max_depth"], p["reachable"]) == (5, 6)
