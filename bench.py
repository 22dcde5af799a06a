#!/usr/bin/env python3
"""Graph-kernel benchmark on one MI355X: triangle count, BFS and PageRank on an R-MAT graph.

Workload (default): R-MAT scale 22, edge factor 16, (a, b, c, d) = (0.57, 0.19, 0.19, 0.05),
seeded, vertex ids permuted, symmetrized, deduplicated, self-loops removed, u32 CSR. One step =
  - triangle_count (symmetric CSR; the count is checked against the committed golden value),
  - BFS from `--sources` seeded random sources of nonzero degree (distances left in HBM),
  - PageRank d = 0.85, `--pr-iterations` iterations, tol 1e-6 (ranks left in HBM).
The graph is uploaded and prepared (degree-ordered layout) before the timed region.

Reported (one JSON line): `value` = whole-step GTEPS with the reference's TEPS conventions
(crates/grafeo-cuda-algos/src/bin/gpu_triangles.rs:127-132, gpu_bfs.rs:78-83,
gpu_pagerank.rs:80-85): (m [TC] + m per BFS source + m x iterations [PageRank]) / step time,
where m = CSR entries. Per-algorithm GTEPS, kernel event times, the roofline of the dominant
kernel (HBM-bound: algorithmic bytes per launch / average launch time, against 8 TB/s) and a CPU
baseline (the C oracle on the host cores, on a bounded sample) go into the same line.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time
from ctypes import byref, c_double, c_int32, c_uint32, c_uint64, c_void_p

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "grafeo-cuda_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS + HBM GB/s (TC, BFS, PageRank) RMAT-22; learned vs fixed sched"
GOLDEN = os.path.join(REPO, "tests", "golden", "rmat_counts.json")
KERNELS = ("tc_bitmap16", "tc_bitmap", "tc_hash", "tc_wedge", "tc_count", "pr_spmv_hot", "pr_spmv", "pr_hub_finish",
           "pr_finalize", "pr_init", "bfs_topdown_edges", "bfs_frontier_degrees", "bfs_bottomup",
           "bfs_queue_to_bitmap", "bfs_bitmap_to_queue", "bfs_init", "bfs_frontier_scan_small", "scan_block_sums",
           "exclusive_scan_small", "scan_apply", "gather_u32", "fill_u32")
# The kernels timed with events in a default run: those the rooflines report. Each timed launch
# adds two event records to the stream (about 2 ms per step if every launch were timed).
ROOFLINE_KERNELS = ("tc_bitmap16", "tc_bitmap", "tc_hash", "tc_wedge", "tc_count", "pr_spmv_hot", "pr_spmv")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pick_sources(offsets: np.ndarray, count: int, seed: int) -> list[int]:
    """`count` distinct seeded random vertices of nonzero degree."""
    deg = np.diff(offsets.astype(np.int64))
    candidates = np.flatnonzero(deg > 0)
    rng = np.random.default_rng(seed)
    return [int(v) for v in rng.choice(candidates, size=min(count, candidates.size), replace=False)]


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--sources", type=int, default=8)
    ap.add_argument("--pr-iterations", type=int, default=20)
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--kernel-events", choices=["on", "all", "off"], default="on",
                    help="time the roofline kernels (on) or every kernel launch (all) of the timed "
                         "steps with HIP events")
    ap.add_argument("--batched", choices=["on", "off"], default="on",
                    help="also time the batched schedules against the one-at-a-time ones: 128-source "
                         "multi-source BFS vs 128 single-source BFS, and the 256-graph triangle-count "
                         "batch (default and autotuned shape) vs one call per graph")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: all host threads (max 16)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes 5 times (the known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings) and exit")
    ap.add_argument("--workload", choices=["step", "scheduler"], default="step",
                    help="step: the R-MAT scale-22 step above; scheduler: the learned schedules "
                         "against the fixed rules on graphs they were not trained on (R-MAT scale 24, "
                         "4096 x 4096 grid)")
    ap.add_argument("--model", default=None, help="learned schedules (default: the committed model)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.workload == "scheduler":
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU", "n_gpus": args.gpus}), flush=True)
            return 2
        return scheduler_workload(args)

    if args.gpus != 1:
        print(
            json.dumps({"error": "this library runs on one GPU (the sharding across GPUs is a design, "
                        "not code), so --gpus must be 1", "n_gpus": args.gpus}),
            flush=True,
        )
        return 2

    import grafeo_cuda
    from grafeo_cuda import _native
    from grafeo_cuda._native import check, lib

    t0 = time.time()
    pairs = _native.gen_rmat(args.scale, args.edge_factor, args.seed)
    offsets, targets = _native.from_edge_pairs(pairs, 1 << args.scale, True)
    del pairs
    n, m = offsets.size - 1, int(targets.size)
    log(f"R-MAT-{args.scale}: n={n} m={m} built in {time.time() - t0:.1f}s")

    dev = grafeo_cuda.GpuDevice(0)
    t0 = time.time()
    g = grafeo_cuda.GpuGraph.from_csr(offsets, targets, dev)
    upload_s = time.time() - t0
    prepare_s = g.prepare()
    m_o, inter, wedges = c_uint64(), c_uint64(), c_uint64()
    check(lib.gc_graph_dag_stats(g._handle, byref(m_o), byref(inter), byref(wedges)))
    log(f"upload {upload_s:.2f}s, prepare {prepare_s:.2f}s, DAG entries {m_o.value}")

    sources = pick_sources(offsets, args.sources, args.seed)
    d_dist, d_rank = c_void_p(), c_void_p()
    check(lib.gc_device_malloc(dev._handle, 4 * n, byref(d_dist)))
    check(lib.gc_device_malloc(dev._handle, 4 * n, byref(d_rank)))

    def step() -> dict:
        times = {}
        tri, secs = c_uint64(), c_double()
        check(lib.gc_triangle_count(g._handle, byref(tri), byref(secs)))
        times["tc"] = secs.value
        bfs_s, reach_total = 0.0, 0
        for s in sources:
            reach, depth = c_uint32(), c_int32()
            check(lib.gc_bfs_device(g._handle, s, d_dist, byref(reach), byref(depth), byref(secs)))
            bfs_s += secs.value
            reach_total += reach.value
        times["bfs"] = bfs_s
        it, conv = c_uint32(), c_int32()
        check(lib.gc_pagerank_device(g._handle, 0.85, args.pr_iterations, 1e-6, d_rank, byref(it), byref(conv), byref(secs)))
        times["pagerank"] = secs.value
        return {"times": times, "triangles": tri.value, "pr_iterations": it.value, "reachable": reach_total}

    for i in range(args.warmup):
        r = step()
        log(f"warmup {i}: {r['times']}")
    # Kernel durations are taken live, with HIP events recorded on the library's stream around
    # every launch of the timed steps.
    if args.kernel_events != "off":
        dev.profile(True, only=ROOFLINE_KERNELS if args.kernel_events == "on" else None)
    t_start = time.perf_counter()
    results = [step() for _ in range(args.steps)]
    elapsed = time.perf_counter() - t_start
    log(f"timed {args.steps} steps in {elapsed:.3f}s")
    kernels = {}
    if args.kernel_events != "off":
        for name in KERNELS:
            ms, count = dev.kernel_time(name)
            if count:
                kernels[name] = {"total_ms": ms / args.steps, "launches": count / args.steps,
                                 "avg_us": 1000.0 * ms / count}
        dev.profile(False)

    iters = results[-1]["pr_iterations"]
    edges_per_step = m * (1 + len(sources) + iters)
    ms_per_step = 1000.0 * elapsed / args.steps
    value = edges_per_step / (elapsed / args.steps) / 1e9
    mean = {k: float(np.mean([r["times"][k] for r in results])) for k in ("tc", "bfs", "pagerank")}
    per_algo = {
        "tc_gteps": m / mean["tc"] / 1e9,
        "bfs_gteps_per_source": m * len(sources) / mean["bfs"] / 1e9,
        "pagerank_gteps": m * iters / mean["pagerank"] / 1e9,
        "tc_ms": 1000 * mean["tc"],
        "bfs_ms_per_source": 1000 * mean["bfs"] / len(sources),
        "pagerank_ms": 1000 * mean["pagerank"],
    }

    # Algorithmic bytes per launch (HBM roofline for every kernel here).
    # PageRank SpMV per iteration: offsets + targets + x read once + y written once.
    spmv_bytes = 4 * (n + 1) + 4 * m + 4 * n + 4 * n
    # Triangle count by list intersection on the DAG (tc_hash, tc_count): offsets, the DAG's
    # entries, and for each DAG edge u->v both out-lists, 4(n+1) + 4 m_o + 4 sum(d+u + d+v).
    tc_bytes = 4 * (n + 1) + 4 * m_o.value + 4 * inter.value
    # Middle-vertex triangle count (tc_wedge): offsets and DAG split (fwd) read once, every
    # CSR entry read once (in-lists and out-lists), plus one 4-byte entry per DAG wedge.
    wedge_bytes = 4 * (n + 1) + 4 * n + 4 * m + 4 * wedges.value
    # BFS per traversal: offsets + targets read once, distances written once.
    bfs_bytes = 4 * (n + 1) + 4 * m + 4 * n
    rooflines = {}
    # HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload
    # (scaled by the readings of a copy of known size), when they have been recorded.
    traffic = load_traffic(f"rmat{args.scale}_ef{args.edge_factor}_seed{args.seed}")

    def add_roofline(name, bytes_per_launch, avg_us):
        a = bytes_per_launch / (avg_us * 1e-6) / 1e9
        t = traffic.get(name)
        rooflines[name] = {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": a / HBM_PEAK_GBS, "traffic": t,
                           "bytes_per_launch": bytes_per_launch,
                           "avg_launch_us": avg_us,
                           # the measured bytes over the same launch time, against the same peak
                           "traffic_frac": (t / (avg_us * 1e-6) / 1e9 / HBM_PEAK_GBS) if t else None}

    # PageRank launches in batches and the launches after convergence return at once: its
    # SpMV time per iteration is the kernel's time per step over the iterations that ran.
    for name in ("pr_spmv", "pr_spmv_hot"):
        if name in kernels:
            add_roofline(name, spmv_bytes, 1000.0 * kernels[name]["total_ms"] / max(1, iters))
            rooflines[name]["note"] = ("avg_launch_us and traffic are per iteration that ran (the "
                                       "launches after convergence return at once)")
            if rooflines[name]["traffic"] is not None:
                # traffic.json averages over all launches of one step: per iteration that ran
                per_iter = rooflines[name]["traffic"] * kernels[name]["launches"] / max(1, iters)
                rooflines[name]["traffic"] = per_iter
                rooflines[name]["traffic_frac"] = (per_iter / (rooflines[name]["avg_launch_us"] * 1e-6)
                                                   / 1e9 / HBM_PEAK_GBS)
    for name in ("tc_bitmap16", "tc_bitmap", "tc_hash", "tc_count"):
        if name in kernels:
            add_roofline(name, tc_bytes, kernels[name]["avg_us"])
    for name in ("tc_bitmap16", "tc_bitmap"):
        if name in rooflines:
            rooflines[name]["note"] = (
                "bytes_per_launch is the merge model (every DAG edge u->v streams both out-lists, "
                "4 bytes an entry); the bitmap kernel reads L(u) once per row into an LDS bitmap "
                "and, with the middle vertices split over the XCDs, serves most L(v) reads from "
                "the L2s (tc_bitmap16 reads the entries of the top 32768 ids as 2-byte offsets), "
                "so it moves fewer bytes (traffic) than the model and frac can exceed 1; "
                "traffic_frac is the measured rate")
    if "tc_wedge" in kernels:
        add_roofline("tc_wedge", wedge_bytes, kernels["tc_wedge"]["avg_us"])
    # BFS over the whole call per source (every level's kernels, launch gaps and the read-back of
    # each level's counters), not kernel time alone.
    add_roofline("bfs_traversal", bfs_bytes, 1000.0 * per_algo["bfs_ms_per_source"])
    rooflines["bfs_traversal"]["note"] = "avg_launch_us is the whole BFS call per source"
    bfs_traffic = bfs_call_traffic(f"rmat{args.scale}_ef{args.edge_factor}_seed{args.seed}")
    if bfs_traffic is not None:
        r = rooflines["bfs_traversal"]
        r["traffic"] = bfs_traffic
        r["traffic_frac"] = bfs_traffic / (r["avg_launch_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS
        r["note"] += ("; traffic sums the measured bytes of the level kernels (bfs_*, without the "
                      "once-per-graph bfs_top_build) over the profiled calls, per call (bfs_init "
                      "dispatches)")
    # The dominant kernel: the one with the most event time in the profiled step.
    dominant = max(kernels, key=lambda k: kernels[k]["total_ms"]) if kernels else None
    roofline = dict(rooflines.get(dominant) or rooflines.get("pr_spmv") or {})
    if roofline:
        roofline["kernel"] = dominant if dominant in rooflines else "pr_spmv"

    parity = {}
    try:
        with open(GOLDEN) as f:
            golden = json.load(f)
        key = f"rmat{args.scale}_ef{args.edge_factor}_seed{args.seed}"
        if key in golden:
            parity["triangles_match_golden"] = results[-1]["triangles"] == golden[key]["triangles"]
    except FileNotFoundError:
        pass

    batched = None
    if args.batched == "on":
        batched = batched_section(dev, g, offsets, n, m, args)

    cpu = None
    if args.cpu_baseline == "on":
        cpu = cpu_baseline(offsets, targets, sources, iters, args.cpu_threads, m)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (TC, BFS); f32 ranks, f64 reductions (PageRank)",
        "data": "synthetic",
        "config": {
            "workload": f"R-MAT scale {args.scale}, edge factor {args.edge_factor}: triangle count + "
                        f"BFS x{len(sources)} sources + PageRank x{iters} iterations per step",
            "num_nodes": n, "csr_entries": m, "dag_entries": m_o.value, "seed": args.seed,
            "bfs_sources": len(sources), "pagerank_iterations": iters,
            "parallelism": "single GPU",
        },
        "per_algorithm": per_algo,
        "roofline": roofline,
        "rooflines": rooflines,
        "kernels": kernels,
        "prepare_seconds": prepare_s,
        "upload_seconds": upload_s,
        "triangles": results[-1]["triangles"],
        "parity": parity,
        "cpu_baseline": cpu,
        "batched": batched,
        "device": dev.name,
    }
    check(lib.gc_device_free(dev._handle, d_dist))
    check(lib.gc_device_free(dev._handle, d_rank))
    print(json.dumps(out), flush=True)
    return 0


def batched_section(dev, g, offsets, n, m, args) -> dict:
    """Batched schedules against the same work done one unit at a time, on the same GPU, same
    inputs, bit-identical results (checked here)."""
    import grafeo_cuda
    from grafeo_cuda import workloads
    from grafeo_cuda._native import check, lib

    out = {}
    # -- 128-source multi-source BFS vs 128 single-source BFS (same sources) --
    sources = pick_sources(offsets, 128, args.seed + 128)
    src = np.asarray(sources, dtype=np.uint32)
    src_p = src.ctypes.data_as(ctypes.POINTER(c_uint32))
    d_ms, d_one = c_void_p(), c_void_p()
    check(lib.gc_device_malloc(dev._handle, 4 * n * len(sources), byref(d_ms)))
    check(lib.gc_device_malloc(dev._handle, 4 * n, byref(d_one)))
    reach = np.zeros(len(sources), np.uint32)
    depth = np.zeros(len(sources), np.int32)
    secs = c_double()

    def run_ms():
        check(lib.gc_msbfs_device(g._handle, src_p, len(sources), d_ms,
                                  reach.ctypes.data_as(ctypes.POINTER(c_uint32)),
                                  depth.ctypes.data_as(ctypes.POINTER(c_int32)), byref(secs)))
        return secs.value

    def run_single():
        total, r, d = 0.0, c_uint32(), c_int32()
        reach1 = []
        for s in sources:
            check(lib.gc_bfs_device(g._handle, s, d_one, byref(r), byref(d), byref(secs)))
            total += secs.value
            reach1.append(r.value)
        return total, reach1

    run_ms()
    ms_t = min(run_ms() for _ in range(3))
    run_single()
    single_t, reach1 = min((run_single() for _ in range(2)), key=lambda x: x[0])
    # row 0 of the batched distances against the single-source distances of source 0
    row = np.empty(n, np.int32)
    one = np.empty(n, np.int32)
    check(lib.gc_device_copy_to_host(dev._handle, row.ctypes.data_as(c_void_p), d_ms, 4 * n))
    rr, dd = c_uint32(), c_int32()
    check(lib.gc_bfs_device(g._handle, sources[0], d_one, byref(rr), byref(dd), byref(secs)))
    check(lib.gc_device_copy_to_host(dev._handle, one.ctypes.data_as(c_void_p), d_one, 4 * n))
    check(lib.gc_device_free(dev._handle, d_ms))
    check(lib.gc_device_free(dev._handle, d_one))
    out["msbfs"] = {
        "sources": len(sources), "ms": 1000 * ms_t, "gteps": m * len(sources) / ms_t / 1e9,
        "single_source_ms": 1000 * single_t, "single_source_gteps": m * len(sources) / single_t / 1e9,
        "speedup": single_t / ms_t,
        "identical": bool(list(reach) == reach1 and np.array_equal(row, one)),
    }
    log(f"msbfs: {out['msbfs']}")

    # -- 256-graph triangle-count batch vs one call per graph --
    t0 = time.time()
    graphs = workloads.uneven_batch(256, 1)
    entries = sum(int(t.size) for _, t in graphs)
    batch = grafeo_cuda.GpuBatch.from_csrs(graphs, dev)
    singles = [grafeo_cuda.GpuGraph.from_csr(o, t, dev) for o, t in graphs]
    for s_g in singles:
        s_g.prepare()
    log(f"batch of 256 graphs ({entries} CSR entries) built in {time.time() - t0:.1f}s")
    expected = None
    try:
        with open(os.path.join(REPO, "tests", "golden", "batch_counts.json")) as f:
            expected = [x["triangles"] for x in json.load(f)["uneven_batch_256_seed1"]["graphs"]]
    except (FileNotFoundError, KeyError, ValueError):
        pass

    def time_batch(reps=10):
        batch.triangle_count()
        best, counts = 1e30, None
        for _ in range(reps):
            r = batch.triangle_count()
            best, counts = min(best, r["time_seconds"]), r["counts"]
        return best, [int(c) for c in counts]

    default_cfg = batch.config
    default_t, counts_default = time_batch()
    tuned = batch.tune()
    tuned_t, counts_tuned = time_batch()

    def time_singles():
        total, counts = 0.0, []
        for s_g in singles:
            r = s_g.triangle_count()
            total += r["time_seconds"]
            counts.append(r["total_triangles"])
        return total, counts

    time_singles()
    single_t, counts_single = min((time_singles() for _ in range(2)), key=lambda x: x[0])
    out["batch_tc"] = {
        "graphs": len(graphs), "csr_entries": entries,
        "default_config": list(default_cfg), "default_ms": 1000 * default_t,
        "tuned_config": list(tuned["config"]), "tuned_ms": 1000 * tuned_t,
        "tuned_gteps": entries / tuned_t / 1e9,
        "one_call_per_graph_ms": 1000 * single_t, "speedup_vs_one_call_per_graph": single_t / tuned_t,
        "speedup_tuned_vs_default": default_t / tuned_t,
        "identical": counts_default == counts_tuned == counts_single,
        "match_golden": None if expected is None else counts_tuned == expected,
    }
    log(f"batch_tc: {out['batch_tc']}")
    return out


def scheduler_workload(args) -> int:
    """Learned schedules against the fixed rules (Beamer's direction switch for BFS; the bitmap
    intersection for triangle counting) on the same GPU and inputs; results must be identical."""
    import grafeo_cuda
    from grafeo_cuda import scheduler
    from grafeo_cuda._native import check, lib

    dev = grafeo_cuda.GpuDevice(0)
    bfs_s, tc_s = scheduler.load(dev, args.model or scheduler.MODEL)
    out = {"metric": "learned vs fixed schedule (BFS per level, TC per degree bucket)",
           "unit": "ms", "higher_is_better": False, "n_gpus": 1, "data": "synthetic", "graphs": {}}
    cases = [("rmat24", lambda: scheduler.rmat(24, 16, 1), 4), ("grid4096", lambda: scheduler.grid(4096, 4096), 0)]
    for name, build, nsrc in cases:
        t0 = time.time()
        off, tgt = build()
        n = off.size - 1
        sources = pick_sources(off, nsrc, args.seed) if nsrc else [n // 2 + 2048, 0]
        g = grafeo_cuda.GpuGraph.from_csr(off, tgt, dev)
        del tgt
        g.prepare()
        log(f"{name}: n={n} m={g.num_edges} ready in {time.time() - t0:.1f}s")
        d_dist = c_void_p()
        check(lib.gc_device_malloc(dev._handle, 4 * n, byref(d_dist)))
        host = np.empty(n, np.int32)

        def bfs_times():
            total, checks = 0.0, []
            for s in sources:
                secs, reach, depth = c_double(), c_uint32(), c_int32()
                check(lib.gc_bfs_device(g._handle, s, d_dist, byref(reach), byref(depth), byref(secs)))
                best = secs.value
                check(lib.gc_bfs_device(g._handle, s, d_dist, byref(reach), byref(depth), byref(secs)))
                best = min(best, secs.value)
                check(lib.gc_device_copy_to_host(dev._handle, host.ctypes.data_as(c_void_p), d_dist, 4 * n))
                checks.append((reach.value, depth.value, hash(host.tobytes())))
                total += best
            return total, checks

        def tc_time():
            g.triangle_count()
            r = [g.triangle_count() for _ in range(2)]
            return min(x["time_seconds"] for x in r), r[0]["total_triangles"]

        g.set_scheduler()
        fixed_bfs, fixed_checks = bfs_times()
        fixed_tc, fixed_tri = tc_time()
        g.set_scheduler(bfs_s, tc_s)
        learned_bfs, learned_checks = bfs_times()
        bfs_schedule = g.last_schedule()
        learned_tc, learned_tri = tc_time()
        tc_schedule = g.last_schedule()
        check(lib.gc_device_free(dev._handle, d_dist))
        out["graphs"][name] = {
            "num_nodes": n, "csr_entries": g.num_edges, "bfs_sources": len(sources),
            "bfs_fixed_ms": 1000 * fixed_bfs, "bfs_learned_ms": 1000 * learned_bfs,
            "bfs_speedup": fixed_bfs / learned_bfs, "bfs_identical": fixed_checks == learned_checks,
            "bfs_last_schedule": {str(c): bfs_schedule.count(c) for c in (0, 1, 2)},
            "tc_fixed_ms": 1000 * fixed_tc, "tc_learned_ms": 1000 * learned_tc,
            "tc_speedup": fixed_tc / learned_tc, "tc_identical": fixed_tri == learned_tri,
            "tc_schedule": tc_schedule, "triangles": fixed_tri,
        }
        log(f"{name}: {out['graphs'][name]}")
        del g
    print(json.dumps(out), flush=True)
    return 0


def calibrate_copy(nbytes: int) -> int:
    """A streaming copy of a known size: reads `nbytes` and writes `nbytes` per launch."""
    import torch

    words = nbytes // 4
    src = torch.arange(words, dtype=torch.int32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel": "torch.neg (int32, elementwise)", "bytes_read": 4 * words,
                      "bytes_written": 4 * words, "launches": 5}), flush=True)
    return 0


def load_traffic(workload_key: str) -> dict:
    """Per-kernel HBM bytes per launch measured with rocprofv3 counters for this workload."""
    try:
        with open(os.path.join(REPO, "profiles", "traffic.json")) as f:
            t = json.load(f)
    except (FileNotFoundError, ValueError):
        return {}
    if t.get("workload") != workload_key:
        return {}
    return {k: v["hbm_bytes_per_launch"] for k, v in t.get("kernels", {}).items()}


def bfs_call_traffic(workload_key: str):
    """HBM bytes of one BFS call: the measured bytes of every level kernel (bfs_*) over the
    profiled calls, divided by the calls (one bfs_init each); None without a measurement."""
    try:
        with open(os.path.join(REPO, "profiles", "traffic.json")) as f:
            t = json.load(f)
    except (FileNotFoundError, ValueError):
        return None
    if t.get("workload") != workload_key:
        return None
    ks = t.get("kernels", {})
    calls = ks.get("bfs_init", {}).get("dispatches", 0)
    if not calls:
        return None
    total = sum(v["hbm_bytes_per_launch"] * v["dispatches"] for k, v in ks.items()
                if k.startswith("bfs_") and k != "bfs_top_build")
    return total / calls


def cpu_baseline(offsets, targets, sources, iters, threads, m) -> dict:
    """The C oracle (a port of the reference's semantics, OpenMP) on the host cores, on a sample:
    one triangle count, one BFS source, two PageRank iterations; scaled to the step's mix."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    threads = threads or min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    t = time.perf_counter()
    oracle.triangle_count(offsets, targets, fast=True)
    tc_s = time.perf_counter() - t
    t = time.perf_counter()
    oracle.bfs(offsets, targets, sources[0])
    bfs_s = time.perf_counter() - t
    t = time.perf_counter()
    oracle.pagerank(offsets, targets, 0.85, 2, 0.0)
    pr_iter_s = (time.perf_counter() - t) / 2
    step_s = tc_s + len(sources) * bfs_s + iters * pr_iter_s
    value = m * (1 + len(sources) + iters) / step_s / 1e9
    log(f"cpu baseline: tc {tc_s:.2f}s bfs {bfs_s:.2f}s pr/it {pr_iter_s:.3f}s on {threads} threads")
    return {
        "value": value,
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": f"1 triangle count, 1 BFS source, 2 PageRank iterations of the same graph, "
                  f"scaled to the step's mix (tc {tc_s:.3f}s, bfs {bfs_s:.3f}s/source, "
                  f"pagerank {pr_iter_s:.4f}s/iteration); host {platform.processor() or platform.machine()}",
    }


if __name__ == "__main__":
    sys.exit(main())
