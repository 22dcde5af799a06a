"""The benchmark command lines end to end on the GPU: each prints the reference's JSON document
with the right answer in `extra`, and --save writes it under results/raw."""

from __future__ import annotations

import json
import os

import pytest

from grafeo_cuda import harness

pytestmark = pytest.mark.gpu


def k5_tsv(tmp_path):
    path = tmp_path / "k5.tsv"
    path.write_text("".join(f"{a}\t{b}\n" for a in range(5) for b in range(a + 1, 5)))
    return str(path)


def run(capsys, argv):
    assert harness.main(argv) == 0
    out = capsys.readouterr().out
    docs, buf, depth = [], "", 0
    for line in out.splitlines():
        buf += line + "\n"
        depth += line.count("{") - line.count("}")
        if depth == 0 and buf.strip():
            docs.append(json.loads(buf))
            buf = ""
    return docs


def test_triangles_symmetric_and_oriented(capsys, tmp_path, device):
    data = k5_tsv(tmp_path)
    (doc,) = run(capsys, ["triangles", "--dataset", data, "--runs", "2"])
    assert doc["algorithm"] == "gpu_triangle_counting" and doc["extra"]["total_triangles"] == 10
    assert doc["vertices"] == 5 and doc["edges"] == 20 and len(doc["runs"]) == 2
    (doc,) = run(capsys, ["triangles", "--dataset", data, "--direct", "--runs", "2"])
    assert doc["algorithm"] == "gpu_triangle_counting_oriented"
    assert doc["extra"] == {"total_triangles": 10, "oriented_edges": 10} and doc["edges"] == 20


def test_bfs_pagerank_ktruss_and_save(capsys, tmp_path, device, monkeypatch):
    data = k5_tsv(tmp_path)
    docs = run(capsys, ["bfs", "--dataset", data, "--sources", "2", "--runs", "1"])
    assert [d["algorithm"] for d in docs] == ["gpu_bfs_source_0", "gpu_bfs_source_1"]
    assert docs[1]["extra"] == {"source_node": 2, "max_depth": 1, "reachable_nodes": 5}
    (doc,) = run(capsys, ["pagerank", "--dataset", data, "--runs", "1"])
    assert doc["extra"]["converged"] is True and doc["extra"]["damping"] == 0.85
    monkeypatch.chdir(tmp_path)
    (doc,) = run(capsys, ["ktruss", "--dataset", data, "--k", "5", "--runs", "1", "--save"])
    assert doc["extra"] == {"k": 5, "truss_edges": 10, "iterations": 1}
    saved = json.loads((tmp_path / "results" / "raw" / "gpu_gpu_ktruss_k5.json").read_text())
    assert saved == doc
    assert os.path.exists(tmp_path / "results" / "raw")
