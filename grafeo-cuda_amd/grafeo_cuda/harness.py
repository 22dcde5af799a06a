"""Benchmark command lines with the reference's options, TEPS conventions and JSON output.

    python3 -m grafeo_cuda.harness triangles --dataset PATH [--runs 5] [--warmup 1] [--direct] [--oriented] [--save]
    python3 -m grafeo_cuda.harness bfs       --dataset PATH [--runs 5] [--warmup 1] [--sources 4] [--direct] [--save]
    python3 -m grafeo_cuda.harness pagerank  --dataset PATH [--damping 0.85] [--max-iterations 20]
                                             [--tolerance 1e-6] [--runs 5] [--warmup 1] [--direct] [--save]
    python3 -m grafeo_cuda.harness ktruss    --dataset PATH [--k 3] [--runs 5] [--warmup 1] [--save]

Each follows its counterpart in crates/grafeo-cuda-algos/src/bin/ (gpu_triangles.rs, gpu_bfs.rs,
gpu_pagerank.rs, gpu_ktruss.rs): warmup calls, then `runs` timed calls of the public entry point,
the best call time as kernel_time_s, TEPS as there (CSR entries / time; times iterations for
PageRank; 2 x oriented entries for the oriented triangle count; per BFS source, with the
harmonic mean over sources on stderr), and one JSON document per result with the keys of
GpuBenchmarkResult (crates/grafeo-cuda-algos/src/harness.rs:13-36), printed to stdout (the
reference's BFS binary only saves its per-source documents; here they are printed too) and, with
--save, written to results/raw/gpu_<algorithm>_<dataset stem>.json.

Loading: the dataset is a TSV edge list read with parse_tsv_to_edges and built into a CSR with
from_edge_pairs (symmetric) or from_edge_pairs_oriented, as the reference's --direct path does
(harness.rs:79-160). `--csr device` (the default) parses the file and builds those same arrays on
the GPU (line breaks and fields per lane, ids numbered by radix sorts, CSR by radix sort,
deduplication, orientation); `--csr host` does both on the host as the reference does. The reference's other path imports the file into a Grafeo store first; Grafeo
is not part of this library, so both paths load the same way here. A dataset may also name a
seeded synthetic graph: rmat:SCALE[:EDGE_FACTOR[:SEED]], er:N:AVG_DEGREE[:SEED] or grid:ROWS:COLS.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

from . import GpuDevice, GpuGraph
from . import _native


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def edges_of(dataset: str):
    """(edge pairs, num_nodes) of a TSV path or a synthetic spec."""
    if dataset.startswith("rmat:"):
        parts = [int(x) for x in dataset.split(":")[1:]]
        scale = parts[0]
        ef = parts[1] if len(parts) > 1 else 16
        seed = parts[2] if len(parts) > 2 else 1
        return _native.gen_rmat(scale, ef, seed), 1 << scale
    if dataset.startswith("er:"):
        parts = dataset.split(":")[1:]
        n, degree = int(parts[0]), float(parts[1])
        seed = int(parts[2]) if len(parts) > 2 else 1
        return _native.gen_erdos_renyi(n, int(n * degree / 2), seed), n
    if dataset.startswith("grid:"):
        rows, cols = (int(x) for x in dataset.split(":")[1:3])
        return _native.gen_grid(rows, cols), rows * cols
    pairs, n = _native.parse_tsv_to_edges(dataset)
    return pairs, n


def load_direct(dataset: str, device: GpuDevice, oriented: bool = False, csr: str = "device"):
    """(GpuGraph, seconds for parse + CSR build + transfer)."""
    start = time.perf_counter()
    synthetic = dataset.startswith(("rmat:", "er:", "grid:"))
    if csr == "device" and not synthetic and not oriented:
        # parsed and built on the GPU: the file's bytes go up, no pairs or CSR come back
        g = GpuGraph.from_tsv(dataset, device)
        total = time.perf_counter() - start
        log(f"  Parsed and built on the GPU: {g.num_nodes} nodes, {g.parsed_pairs} edges, {g.num_edges} CSR "
            f"entries in {total:.4f}s from {dataset}")
        return g, total
    if csr == "device" and not synthetic:
        pairs, n = _native.parse_tsv_to_edges_device(device._handle, dataset)
    else:
        pairs, n = edges_of(dataset)
    log(f"  Parsed {n} nodes, {len(pairs)} edges in {time.perf_counter() - start:.2f}s from {dataset}")
    t = time.perf_counter()
    if csr == "device" and not oriented:
        # built on the GPU and kept there: the pairs go up, no CSR comes back
        g = GpuGraph.from_edge_pairs(pairs, n, device)
        log(f"  CSR build on the GPU (with the pairs' transfer): {time.perf_counter() - t:.4f}s "
            f"({n} nodes, {g.num_edges} CSR entries)")
    else:
        if csr == "device":
            off, tgt = _native.from_edge_pairs_oriented_device(device._handle, pairs, n)
        elif oriented:
            off, tgt = _native.from_edge_pairs_oriented(pairs, n)
        else:
            off, tgt = _native.from_edge_pairs(pairs, n, True)
        log(f"  CSR build ({csr}): {time.perf_counter() - t:.4f}s ({n} nodes, {len(tgt)} CSR entries)")
        t = time.perf_counter()
        g = GpuGraph.from_csr(off, tgt, device)
        log(f"  GPU transfer: {time.perf_counter() - t:.4f}s")
    total = time.perf_counter() - start
    log(f"  Total (parse+CSR+transfer): {total:.4f}s  [{g.num_nodes} vertices, {g.num_edges} CSR entries]")
    return g, total


def platform_info(device: GpuDevice) -> str:
    """OS / architecture / host name (COMPUTERNAME or HOSTNAME, else "unknown") / GPU name, as
    harness.rs:162-176 formats it."""
    host = os.environ.get("COMPUTERNAME") or os.environ.get("HOSTNAME") or "unknown"
    return f"{sys.platform} / {platform.machine()} / {host} / GPU: {device.name}"


def result(dataset, algorithm, vertices, edges, transfer, times, teps, device, extra) -> dict:
    best = min(times) if times else 0.0
    out = {
        "dataset": dataset,
        "algorithm": algorithm,
        "vertices": int(vertices),
        "edges": int(edges),
        "time_s": transfer + best,
        "transfer_time_s": transfer,
        "kernel_time_s": best,
        "teps": teps,
        "runs": list(times),
        "platform": platform_info(device),
        "gpu_device": device.name,
    }
    if extra is not None:
        out["extra"] = extra
    return out


def save(res: dict) -> str:
    directory = os.path.join("results", "raw")
    os.makedirs(directory, exist_ok=True)
    stem = os.path.splitext(os.path.basename(res["dataset"]))[0] or "unknown"
    path = os.path.join(directory, f"gpu_{res['algorithm']}_{stem.replace(':', '_')}.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=2)
    return path


def emit(res: dict, do_save: bool) -> None:
    if do_save:
        log(f"  Saved to {save(res)}")
    print(json.dumps(res, indent=2), flush=True)


def timed(call, warmup: int, runs: int):
    for _ in range(warmup):
        call()
    times, last = [], None
    for _ in range(runs):
        last = call()
        times.append(last["time_seconds"])
    return times, last


def run_triangles(args, device) -> list[dict]:
    oriented = args.oriented or args.direct
    g, transfer = load_direct(args.dataset, device, oriented=oriented, csr=args.csr)
    call = g.triangle_count_oriented if oriented else g.triangle_count
    times, last = timed(call, args.warmup, args.runs)
    best = min(times)
    edges = 2 * g.num_edges if oriented else g.num_edges
    teps = edges / best if best > 0 else 0.0
    name = "gpu_triangle_counting_oriented" if oriented else "gpu_triangle_counting"
    log(f"\n  [{name}] on {args.dataset}\n    Triangles: {last['total_triangles']}\n"
        f"    Transfer:  {transfer:.6f}s\n    Kernel:    {best:.6f}s\n    TEPS:      {teps:.0f}")
    extra = {"total_triangles": last["total_triangles"]}
    if oriented:
        extra["oriented_edges"] = g.num_edges
    res = result(args.dataset, name, g.num_nodes, edges, transfer, times, teps, device, extra)
    emit(res, args.save)
    return [res]


def run_bfs(args, device) -> list[dict]:
    g, transfer = load_direct(args.dataset, device, csr=args.csr)
    n = g.num_nodes
    if n == 0:
        log("ERROR: Graph is empty")
        sys.exit(1)
    step = max(1, n // max(1, args.sources))
    sources = [i * step for i in range(args.sources)]
    out, all_teps = [], []
    for i, src in enumerate(sources):
        log(f"\n  BFS source {i + 1}/{len(sources)}: node {src}")
        times, last = timed(lambda: g.bfs_numpy(src), args.warmup, args.runs)
        best = min(times)
        teps = g.num_edges / best if best > 0 else 0.0
        log(f"    Depth: {last['max_depth']}, Reachable: {last['reachable']}\n    Transfer: {transfer:.6f}s\n"
            f"    Kernel:   {best:.6f}s\n    TEPS:     {teps:.0f}")
        all_teps.append(teps)
        res = result(args.dataset, f"gpu_bfs_source_{i}", n, g.num_edges, transfer, times, teps, device,
                     {"source_node": src, "max_depth": last["max_depth"], "reachable_nodes": last["reachable"]})
        if args.save:
            save(res)
        out.append(res)
    positive = [t for t in all_teps if t > 0]
    if positive:
        harmonic = len(positive) / sum(1.0 / t for t in positive)
        log(f"\n  Harmonic mean TEPS across {len(sources)} sources: {harmonic:.0f}")
    for res in out:
        print(json.dumps(res, indent=2), flush=True)
    return out


def run_pagerank(args, device) -> list[dict]:
    g, transfer = load_direct(args.dataset, device, csr=args.csr)
    call = lambda: g.pagerank_numpy(args.damping, args.max_iterations, args.tolerance)  # noqa: E731
    times, last = timed(call, args.warmup, args.runs)
    best = min(times)
    teps = g.num_edges * last["iterations"] / best if best > 0 else 0.0
    order = np.argsort(-last["ranks"], kind="stable")[:5]
    top5 = [(int(i), float(last["ranks"][i])) for i in order]
    log(f"\n  [gpu_pagerank] on {args.dataset}\n    Iterations: {last['iterations']}, Converged: "
        f"{str(last['converged']).lower()}\n    Transfer:   {transfer:.6f}s\n    Kernel:     {best:.6f}s\n"
        f"    TEPS:       {teps:.0f}\n    Top 5:      {top5}")
    res = result(args.dataset, "gpu_pagerank", g.num_nodes, g.num_edges, transfer, times, teps, device,
                 {"iterations": last["iterations"], "converged": last["converged"], "damping": args.damping})
    emit(res, args.save)
    return [res]


def run_ktruss(args, device) -> list[dict]:
    g, transfer = load_direct(args.dataset, device, csr=args.csr)
    times, last = timed(lambda: g.k_truss_numpy(args.k), args.warmup, args.runs)
    best = min(times)
    teps = g.num_edges / best if best > 0 else 0.0
    log(f"\n  [gpu_ktruss] on {args.dataset} (k={args.k})\n    Edges in {args.k}-truss: {len(last['edges'])}\n"
        f"    Iterations: {last['iterations']}\n    Transfer:   {transfer:.6f}s\n    Kernel:     {best:.6f}s\n"
        f"    TEPS:       {teps:.0f}")
    res = result(args.dataset, "gpu_ktruss", g.num_nodes, g.num_edges, transfer, times, teps, device,
                 {"k": args.k, "truss_edges": int(len(last["edges"])), "iterations": last["iterations"]})
    emit(res, args.save)
    return [res]


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="python3 -m grafeo_cuda.harness", description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)

    def common(p, direct=True):
        p.add_argument("--dataset", required=True, help="TSV edge list, or rmat:/er:/grid: spec")
        p.add_argument("--runs", type=int, default=5)
        p.add_argument("--warmup", type=int, default=1)
        if direct:
            p.add_argument("--direct", action="store_true", help="TSV -> CSR -> GPU (the only path here)")
        p.add_argument("--save", action="store_true", help="write results/raw/gpu_<algorithm>_<stem>.json")
        p.add_argument("--csr", choices=["device", "host"], default="device",
                       help="build the CSR on the GPU (default) or on the host as the reference does")

    t = sub.add_parser("triangles", help="gpu_triangles")
    common(t)
    t.add_argument("--oriented", action="store_true", help="degree-oriented CSR (default with --direct)")
    b = sub.add_parser("bfs", help="gpu_bfs")
    common(b)
    b.add_argument("--sources", type=int, default=4, help="evenly spaced BFS sources")
    p = sub.add_parser("pagerank", help="gpu_pagerank")
    common(p)
    p.add_argument("--damping", type=float, default=0.85)
    p.add_argument("--max-iterations", type=int, default=20)
    p.add_argument("--tolerance", type=float, default=1e-6)
    k = sub.add_parser("ktruss", help="gpu_ktruss")
    common(k, direct=False)
    k.add_argument("--k", type=int, default=3)
    return ap


def main(argv=None) -> int:
    args = parser().parse_args(argv)
    device = GpuDevice(0)
    {"triangles": run_triangles, "bfs": run_bfs, "pagerank": run_pagerank, "ktruss": run_ktruss}[args.cmd](args, device)
    return 0


if __name__ == "__main__":
    sys.exit(main())
