"""CSR construction on the GPU: the same arrays as the host builder (from_edge_pairs,
crates/grafeo-cuda-core/src/csr_export.rs:202-239) for symmetric and directed inputs with
self-loops, duplicates and isolated nodes, at small and full sizes, and the same error for
out-of-range ids."""

from __future__ import annotations

import json
import os
import time

import numpy as np
import pytest

import grafeo_cuda
from grafeo_cuda import _native
from graphs import csr_sha256, random_graph

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rmat_counts.json")


def same(device, pairs, n, symmetric):
    host = _native.from_edge_pairs(pairs, n, symmetric)
    dev = _native.from_edge_pairs_device(device._handle, pairs, n, symmetric)
    assert np.array_equal(host[0], dev[0]) and np.array_equal(host[1], dev[1])
    return host


def test_small_inputs(device):
    assert same(device, [], 0, True)[0].tolist() == [0]
    assert same(device, [], 5, True)[0].tolist() == [0] * 6
    off, tgt = same(device, [(0, 0), (0, 0)], 1, True)
    assert off.tolist() == [0, 1] and tgt.tolist() == [0]
    rng = np.random.default_rng(5)
    for _ in range(60):
        n, edges = random_graph(rng, max_nodes=40, max_edges=200)
        for symmetric in (True, False):
            same(device, edges, n, symmetric)


def test_large_inputs(device):
    rng = np.random.default_rng(6)
    for n, count in ((1, 1000), (2, 5000), (300, 100000), (70000, 1_000_000), (1 << 20, 3_000_000)):
        pairs = rng.integers(0, n, size=(count, 2), dtype=np.int64).astype(np.uint32)
        for symmetric in (True, False):
            same(device, pairs, n, symmetric)


def test_rmat20_matches_the_golden_csr_and_counts(device):
    with open(GOLDEN) as f:
        golden = json.load(f)["rmat20_ef16_seed1"]
    pairs = _native.gen_rmat(20, 16, 1)
    t = time.perf_counter()
    host = _native.from_edge_pairs(pairs, 1 << 20, True)
    host_s = time.perf_counter() - t
    t = time.perf_counter()
    dev = _native.from_edge_pairs_device(device._handle, pairs, 1 << 20, True)
    dev_s = time.perf_counter() - t
    assert csr_sha256(*dev) == golden["csr_sha256"] == csr_sha256(*host)
    print(f"RMAT-20 CSR build: host {host_s:.3f}s, device {dev_s:.3f}s (with transfers)")
    g = grafeo_cuda.GpuGraph.from_edge_pairs(pairs, 1 << 20, device)
    assert g.num_edges == len(host[1])
    assert g.triangle_count()["total_triangles"] == golden["triangles"]


def test_out_of_range_ids_are_rejected_like_the_host_builder(device):
    pairs = [(0, 1), (2, 7), (9, 1)]
    with pytest.raises(_native.GcError) as host_err:
        _native.from_edge_pairs(pairs, 5, True)
    with pytest.raises(_native.GcError) as dev_err:
        _native.from_edge_pairs_device(device._handle, pairs, 5, True)
    assert str(host_err.value) == str(dev_err.value)
    assert dev_err.value.kind == "InvalidCsr"
    with pytest.raises(RuntimeError, match="node id >= num_nodes"):
        grafeo_cuda.GpuGraph.from_edge_pairs(pairs, 5, device)


def dag_stats(g):
    from ctypes import byref, c_uint64

    a, b, c = c_uint64(), c_uint64(), c_uint64()
    _native.check(_native.lib.gc_graph_dag_stats(g._handle, byref(a), byref(b), byref(c)))
    return a.value, b.value, c.value


def test_device_preparation_builds_the_host_layout(device, monkeypatch):
    """The degree-ordered layout built on the device equals the host restatement's array for
    array (offsets, targets, perm, fwd), with the same symmetry flag and DAG statistics."""
    from graphs import catalog, rmat_graph, undirected

    cases = [(name, off, tgt) for name, off, tgt in catalog()]
    rng = np.random.default_rng(77)
    for i in range(30):
        n, edges = random_graph(rng, max_nodes=30, max_edges=120)
        cases.append((f"random #{i}", *undirected(n, edges)))
        cases.append((f"directed #{i}", *_native.from_edge_pairs(edges, n, False)))
    cases.append(("RMAT-14", *rmat_graph(14, 16, seed=2)))
    cases.append(("RMAT-20", *rmat_graph(20, 16, seed=1)))
    for name, off, tgt in cases:
        monkeypatch.delenv("GC_PREPARE", raising=False)
        gd = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
        t = time.perf_counter()
        dev_layout = gd.relabelled_numpy()
        dev_s = time.perf_counter() - t
        monkeypatch.setenv("GC_PREPARE", "host")
        gh = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
        t = time.perf_counter()
        host_layout = gh.relabelled_numpy()
        host_s = time.perf_counter() - t
        for key in ("offsets", "targets", "perm", "fwd"):
            assert np.array_equal(dev_layout[key], host_layout[key]), (name, key)
        assert gd.is_symmetric() == gh.is_symmetric(), name
        assert dag_stats(gd) == dag_stats(gh), name
        if name.startswith("RMAT-20"):
            print(f"RMAT-20 preparation: device {dev_s:.3f}s, host {host_s:.3f}s")
    monkeypatch.delenv("GC_PREPARE", raising=False)


# ---- oriented builder (from_edge_pairs_oriented, csr_export.rs:299-350) ----------


def same_oriented(device, pairs, n, with_oracle=False):
    host = _native.from_edge_pairs_oriented(pairs, n)
    dev = _native.from_edge_pairs_oriented_device(device._handle, pairs, n)
    assert np.array_equal(host[0], dev[0]) and np.array_equal(host[1], dev[1])
    if with_oracle:
        import oracle

        ref = oracle.from_edge_pairs_oriented(pairs, n)
        assert np.array_equal(ref[0], dev[0]) and np.array_equal(ref[1], dev[1])
    return dev


def test_oriented_small_inputs(device):
    assert same_oriented(device, [], 0)[0].tolist() == [0]
    assert same_oriented(device, [], 5)[0].tolist() == [0] * 6
    # loops only: no edge survives
    off, tgt = same_oriented(device, [(0, 0), (3, 3), (3, 3)], 4, with_oracle=True)
    assert off.tolist() == [0] * 5 and tgt.size == 0
    # a star: every edge leaves its leaf (degree 1) for the centre (degree 3); a duplicate
    # in both directions counts once
    off, tgt = same_oriented(device, [(0, 1), (2, 0), (0, 3), (1, 0)], 4, with_oracle=True)
    assert off.tolist() == [0, 0, 1, 2, 3] and tgt.tolist() == [0, 0, 0]
    # equal degrees (a triangle): from the lower id
    off, tgt = same_oriented(device, [(2, 1), (1, 0), (0, 2)], 3, with_oracle=True)
    assert off.tolist() == [0, 2, 3, 3] and tgt.tolist() == [1, 2, 2]
    rng = np.random.default_rng(15)
    for _ in range(60):
        n, edges = random_graph(rng, max_nodes=40, max_edges=200)
        same_oriented(device, edges, n, with_oracle=True)


def test_oriented_large_inputs(device):
    rng = np.random.default_rng(16)
    for n, count in ((1, 1000), (2, 5000), (300, 100000), (70000, 1_000_000), (1 << 20, 3_000_000)):
        pairs = rng.integers(0, n, size=(count, 2), dtype=np.int64).astype(np.uint32)
        same_oriented(device, pairs, n)


def test_oriented_rmat20_triangles(device):
    """The oriented R-MAT 20 CSR built on the GPU equals the host builder's, and its oriented
    triangle count is the golden count."""
    with open(GOLDEN) as f:
        golden = json.load(f)["rmat20_ef16_seed1"]
    pairs = _native.gen_rmat(20, 16, 1)
    off, tgt = same_oriented(device, pairs, 1 << 20)
    g = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
    assert g.triangle_count_oriented()["total_triangles"] == golden["triangles"]


def test_oriented_out_of_range_ids_are_rejected_like_the_host_builder(device):
    pairs = [(0, 1), (2, 7), (9, 1)]
    with pytest.raises(_native.GcError) as host_err:
        _native.from_edge_pairs_oriented(pairs, 5)
    with pytest.raises(_native.GcError) as dev_err:
        _native.from_edge_pairs_oriented_device(device._handle, pairs, 5)
    assert str(host_err.value) == str(dev_err.value)
    assert dev_err.value.kind == "InvalidCsr"
