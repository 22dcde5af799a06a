"""The batched workload generator is deterministic: it reproduces the committed fixture (sizes,
CSR hashes and oracle triangle counts of all 256 graphs)."""

from __future__ import annotations

import json
import os

import oracle
from graphs import csr_sha256
from grafeo_cuda import workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batch_counts.json")


def test_uneven_batch_matches_fixture():
    with open(GOLDEN) as f:
        fixture = json.load(f)["uneven_batch_256_seed1"]
    graphs = workloads.uneven_batch(256, 1)
    assert len(graphs) == len(fixture["graphs"]) == 256
    for i, ((off, tgt), g) in enumerate(zip(graphs, fixture["graphs"])):
        assert len(off) - 1 == g["num_nodes"] and len(tgt) == g["csr_entries"], i
        assert 1024 <= g["num_nodes"] <= 65536, i
        assert csr_sha256(off, tgt) == g["csr_sha256"], i
    # the oracle's counts of a spread of graphs (all 256 are checked on the GPU)
    for i in range(0, 256, 37):
        off, tgt = graphs[i]
        assert oracle.triangle_count(off, tgt, fast=True) == fixture["graphs"][i]["triangles"], i
