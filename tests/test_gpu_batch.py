"""Batched triangle counting on the GPU: every graph's count is bit-exact against the oracle (and
the committed fixture for the 256-graph workload) whatever the launch shape; a batch replays
its captured graph repeatedly with identical results; bad inputs are rejected with the graph's
index before any device work."""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

import grafeo_cuda
import oracle
from grafeo_cuda import workloads
from graphs import catalog, random_graph, undirected

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batch_counts.json")

SHAPES = [(0, 256, 1, 8), (2, 1, 1, 1), (4, 64, 16, 4), (32, 2048, 4, 8), (128, 8192, 1, 16)]


def small_graphs():
    graphs = [(off, tgt) for _, off, tgt in catalog()]
    rng = np.random.default_rng(99)
    for _ in range(60):
        n, edges = random_graph(rng, max_nodes=40, max_edges=200)
        graphs.append(undirected(n, edges))
    return graphs


def test_small_graphs_every_shape(device):
    graphs = small_graphs()
    expected = np.array([oracle.triangle_count(o, t) for o, t in graphs], dtype=np.uint64)
    b = grafeo_cuda.GpuBatch.from_csrs(graphs, device)
    assert b.num_graphs == len(graphs)
    assert np.array_equal(b.triangle_count()["counts"], expected)
    for shape in SHAPES:
        b.set_config(shape)
        assert b.config == shape
        assert np.array_equal(b.triangle_count()["counts"], expected), shape


def test_empty_batch_and_edgeless_graphs(device):
    b = grafeo_cuda.GpuBatch.from_csrs([], device)
    assert b.num_graphs == 0 and b.triangle_count()["counts"].size == 0
    b = grafeo_cuda.GpuBatch.from_csrs([([0], []), ([0, 0, 0], []), ([0, 1, 1], [0])], device)
    assert list(b.triangle_count()["counts"]) == [0, 0, 0]


def test_rejects_bad_graphs_with_their_index(device):
    k3 = undirected(3, [(0, 1), (1, 2), (0, 2)])
    with pytest.raises(RuntimeError, match=r"graph 1"):
        grafeo_cuda.GpuBatch.from_csrs([k3, ([0, 1], [5])], device)
    with pytest.raises(RuntimeError, match=r"symmetric CSRs \(graph 2\)"):
        grafeo_cuda.GpuBatch.from_csrs([k3, k3, ([0, 1, 1], [1])], device)
    b = grafeo_cuda.GpuBatch.from_csrs([k3], device)
    with pytest.raises(RuntimeError, match="invalid batch config"):
        b.set_config((8, 0, 1, 8))


def test_uneven_256_graph_batch(device):
    with open(GOLDEN) as f:
        fixture = json.load(f)["uneven_batch_256_seed1"]
    graphs = workloads.uneven_batch(256, 1)
    expected = np.array([g["triangles"] for g in fixture["graphs"]], dtype=np.uint64)
    live = np.array([oracle.triangle_count(o, t, fast=True) for o, t in graphs[:: 16]], dtype=np.uint64)
    assert np.array_equal(live, expected[:: 16])
    b = grafeo_cuda.GpuBatch.from_csrs(graphs, device)
    first = b.triangle_count()
    assert np.array_equal(first["counts"], expected)
    tuned = b.tune()
    assert b.config == tuned["config"]
    print(f"batch of 256: default {first['time_seconds'] * 1e3:.2f} ms per call, tuned {tuned}")
    # replays of the captured graph, with other device work in between
    g = grafeo_cuda.GpuGraph.from_csr(*graphs[3], device)
    for _ in range(4):
        assert np.array_equal(b.triangle_count()["counts"], expected)
        assert g.triangle_count()["total_triangles"] == expected[3]
