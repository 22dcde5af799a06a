"""The benchmark command lines: the reference binaries' options and defaults
(crates/grafeo-cuda-algos/src/bin/*.rs) and the GpuBenchmarkResult JSON keys
(crates/grafeo-cuda-algos/src/harness.rs:13-36). No GPU needed."""

from __future__ import annotations

from grafeo_cuda import harness


def test_defaults_match_the_reference_binaries():
    p = harness.parser()
    t = p.parse_args(["triangles", "--dataset", "x.tsv"])
    assert (t.runs, t.warmup, t.direct, t.oriented, t.save) == (5, 1, False, False, False)
    b = p.parse_args(["bfs", "--dataset", "x.tsv"])
    assert (b.runs, b.warmup, b.sources) == (5, 1, 4)
    r = p.parse_args(["pagerank", "--dataset", "x.tsv"])
    assert (r.damping, r.max_iterations, r.tolerance, r.runs, r.warmup) == (0.85, 20, 1e-6, 5, 1)
    k = p.parse_args(["ktruss", "--dataset", "x.tsv"])
    assert (k.k, k.runs, k.warmup) == (3, 5, 1)


class FakeDevice:
    name = "AMD Instinct MI355X"


def test_result_has_the_reference_keys():
    res = harness.result("data/g.tsv", "gpu_pagerank", 10, 40, 0.5, [0.3, 0.2, 0.25], 123.0, FakeDevice(),
                         {"iterations": 20})
    assert list(res) == ["dataset", "algorithm", "vertices", "edges", "time_s", "transfer_time_s",
                         "kernel_time_s", "teps", "runs", "platform", "gpu_device", "extra"]
    assert res["kernel_time_s"] == 0.2 and res["time_s"] == 0.7
    assert res["platform"].endswith("/ GPU: AMD Instinct MI355X")
    assert "extra" not in harness.result("g", "a", 1, 1, 0.0, [1.0], 1.0, FakeDevice(), None)


def test_synthetic_dataset_specs():
    pairs, n = harness.edges_of("grid:3:4")
    assert n == 12 and len(pairs) == 3 * 3 + 2 * 4
    pairs, n = harness.edges_of("rmat:6:4:2")
    assert n == 64 and len(pairs) <= 4 * 64
    pairs, n = harness.edges_of("er:100:8:3")
    assert n == 100 and len(pairs) <= 400


def test_tsv_datasets_use_the_reference_parser(tmp_path):
    path = tmp_path / "k4.tsv"
    path.write_text("# K4\n10\t20\n10\t30\n10\t40\n20\t30\n20\t40\n30\t40\n")
    pairs, n = harness.edges_of(str(path))
    assert n == 4 and len(pairs) == 6
