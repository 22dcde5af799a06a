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
where m = CSR entries. Beside it, outside the timed steps:
  - per-algorithm GTEPS, kernel event times and the roofline of every reported kernel (HBM-bound:
    the stated model's bytes per launch / average launch time, against 8 TB/s), the dominant
    kernel's as `roofline`;
  - the reference's own convention (host outputs, best of 5 calls after 1 warmup, BFS TEPS as
    the harmonic mean over sources) and a Graph500-style BFS GTEPS;
  - the batched schedules against one call at a time (128-source multi-source BFS, the
    256-graph triangle-count batch), by call time and by kernel time;
  - the learned schedules against the fixed rules and every forced choice on R-MAT 22, R-MAT 24
    and the 4096 x 4096 grid (>= 5 runs each, 95% confidence intervals);
  - the CPU baseline (the C oracle on the host's threads and on one thread, on a bounded sample)
    and the PageRank precision study at R-MAT 22 (50 iterations, tol 1e-7, against the FP64
    oracle).
"""

from __future__ import annotations

import argparse
import ctypes
import hashlib
import json
import math
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
KERNEL_FAMILIES = {
    "tc": ("tc_bitmap16", "tc_mid16", "tc_mid32", "tc_bitmap16_w4", "tc_bitmap", "tc_hash", "tc_wedge", "tc_count",
           "tc_mid_rows", "tc_mid_fill", "tc_mid_records", "tc_mid32_records", "tc_mid_rowwork"),
    "pagerank": ("pr_spmv_slice", "pr_slice_finish", "pr_spmv_bin", "pr_bin_scatter", "pr_bin_accum",
                 "pr_spmv_merge", "pr_merge_finish",
                 "pr_spmv_push", "pr_push_finish", "pr_spmv_hot", "pr_spmv",
                 "pr_hub_finish", "pr_finalize", "pr_init", "pr_push_split", "pr_merge_starts", "pr_merge_chunks"),
    "bfs": ("bfs_init", "bfs_topdown_edges", "bfs_frontier_degrees", "bfs_frontier_scan_small", "bfs_bottomup",
            "bfs_queue_to_bitmap", "bfs_bitmap_to_queue", "bfs_publish", "bfs_chain_expand", "bfs_chain_advance",
            "bfs_top_build", "bfs_chain"),
    "msbfs": ("ms_seed", "ms_seed_levels", "ms_topdown", "ms_bottomup_thread", "ms_bottomup_wave",
              "ms_finalize", "ms_write_levels", "ms_levels_to_dist", "ms_write_distances", "ms_summarize"),
    "ktruss": ("kt_bitmap_support", "kt_mid_support", "kt_mid_records", "kt_wedge_support", "kt_truss_frontier", "kt_truss_peel", "kt_truss_commit", "kt_truss_decrement",
               "kt_truss_values", "kt_filter", "kt_support_direct", "kt_reverse_index", "kt_canonical_index",
               "kt_canonical_reverse", "kt_emit_flags_direct", "kt_emit_flags_state"),
    "batch": ("bt_count",),
    "util": ("scan_block_sums", "exclusive_scan_small", "scan_apply", "gather_u32", "fill_u32", "compact_count",
             "compact_emit_pairs", "csr_forward_start"),
}
KERNELS = tuple(k for ks in KERNEL_FAMILIES.values() for k in ks)
# The kernels timed with events in a default run: those the rooflines report. Each timed launch
# adds two event records to the stream (about 2 ms per step if every launch were timed).
ROOFLINE_KERNELS = ("tc_bitmap16", "tc_mid16", "tc_mid32", "tc_bitmap", "tc_hash", "tc_wedge", "tc_count", "pr_spmv_slice", "pr_slice_finish",
                    "pr_spmv_bin", "pr_merge_finish", "pr_bin_scatter", "pr_bin_accum", "pr_spmv_merge",
                    "pr_spmv_push", "pr_spmv_hot", "pr_spmv")
# The binned SpMV (the default on symmetric graphs), four kernels per iteration: the scatter and
# the accumulation of the cold entries, the pull over the hot stream (sliced, the default, or
# chunked: GC_PR_KERNEL=bin) and the finishing of the hub rows and of the rows the pull leaves
# partial.
PR_BIN_SETS = (("pr_bin_scatter", "pr_bin_accum", "pr_spmv_slice", "pr_slice_finish"),
               ("pr_bin_scatter", "pr_bin_accum", "pr_spmv_bin", "pr_merge_finish"))
PR_LAYOUT_KEYS = ("push_count", "hot_count", "stage", "bin_rows", "bin_values", "cold_entries", "values",
                  "slot_rows", "hot_entries", "chunks", "subs", "bins", "runs", "pieces", "piece_entries",
                  "pull_rows", "cold_sources", "build_ns", "grid", "sliced", "slices", "slice_entries",
                  "slice_items", "long_slices", "partial_slots")


def pagerank_layout(g) -> dict:
    """Sizes of the binned layout the graph's last PageRank call used (gc_pagerank_layout)."""
    from grafeo_cuda._native import check, lib

    vals = (c_uint64 * len(PR_LAYOUT_KEYS))()
    check(lib.gc_pagerank_layout(g._handle, vals, len(PR_LAYOUT_KEYS)))
    return dict(zip(PR_LAYOUT_KEYS, (int(v) for v in vals)))


def pagerank_bin_models(lay: dict, zero_rows: int) -> dict:
    """Stored-width bytes per launch of the binned SpMV's kernels, from the layout's sizes:
    - pr_bin_scatter: the 16-bit slots (64 per slot row), a 4-byte group id per block of eight
      slot rows, each cold source's contribution (4 bytes), 8 bytes per run, each cold value
      written (4 bytes);
    - pr_bin_accum: each value and its 16-bit row index (6 bytes) and each pull row's cold sum
      written (8 bytes);
    - pr_spmv_slice (sliced layout): the sliced 16-bit entries with their padding, 16 bytes per
      work item; per pull row its contribution, input rank, cold sum, offset (20 bytes) and rank
      and contribution written (8), less the rows of the wide slices, which write a partial sum
      (8 bytes per row and piece) instead; the hub pieces' 16-bit entries (512 per piece); per workgroup the hot
      set read (4 bytes per id) and the push accumulators written (8 per hub); the rows of
      degree 0 (rank read, rank and contribution written: 12 bytes);
    - pr_slice_finish: the push accumulators of every workgroup read (8 bytes per hub each) and
      the hub rows finished (offsets, input rank, rank and contribution: 16 bytes); the wide
      slices' rows: their partials read, cold sum, offsets and input rank (16 bytes) and rank and
      contribution written (8);
    - pr_spmv_bin (chunked stream): the 16-bit stream in whole 512-entry chunks with a start bit
      per entry, the row data as pr_spmv_slice, 20 bytes of head, tail and tail row per chunk;
    - pr_merge_finish: the push accumulators and hub rows as pr_slice_finish, the chunks' heads,
      tails and tail rows read (20 bytes per chunk) and a row finished per tail (24 bytes)."""
    stream = 512 * lay["chunks"]
    per_wg = lay["grid"] * (4 * lay["hot_count"] + 8 * lay["push_count"])
    hubs = lay["grid"] * 8 * lay["push_count"] + 16 * lay["push_count"]
    wide_rows = 64 * lay["long_slices"]
    return {
        "pr_bin_scatter": 2 * 64 * lay["slot_rows"] + 4 * (lay["slot_rows"] // 8) + 4 * lay["cold_sources"]
        + 8 * lay["runs"] + 4 * lay["cold_entries"],
        "pr_bin_accum": 6 * lay["values"] + 8 * lay["pull_rows"],
        "pr_spmv_slice": 2 * lay["slice_entries"] + 16 * lay["slice_items"] + 28 * lay["pull_rows"]
        - 8 * wide_rows + 512 * lay["partial_slots"] + 1024 * lay["pieces"] + per_wg + 12 * zero_rows,
        "pr_slice_finish": hubs + 512 * lay["partial_slots"] + 24 * wide_rows,
        "pr_spmv_bin": 2 * stream + stream // 8 + 28 * lay["pull_rows"] + 20 * lay["chunks"]
        + 4 * lay["piece_entries"] + per_wg + 12 * zero_rows,
        "pr_merge_finish": hubs + 20 * lay["chunks"] + 24 * min(lay["chunks"], lay["pull_rows"]),
    }
# two-sided 95% quantiles of Student's t by degrees of freedom
T975 = {1: 12.706, 2: 4.303, 3: 3.182, 4: 2.776, 5: 2.571, 6: 2.447, 7: 2.365, 8: 2.306, 9: 2.262, 10: 2.228}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pick_sources(offsets: np.ndarray, count: int, seed: int) -> list[int]:
    """`count` distinct seeded random vertices of nonzero degree."""
    deg = np.diff(offsets.astype(np.int64))
    candidates = np.flatnonzero(deg > 0)
    rng = np.random.default_rng(seed)
    return [int(v) for v in rng.choice(candidates, size=min(count, candidates.size), replace=False)]


def mean_ci(xs) -> dict:
    """Mean, sample standard deviation and the half-width of the 95% confidence interval."""
    a = np.asarray(xs, dtype=np.float64)
    n = a.size
    m = float(a.mean())
    sd = float(a.std(ddof=1)) if n > 1 else 0.0
    half = T975.get(n - 1, 1.96) * sd / math.sqrt(n) if n > 1 else 0.0
    return {"mean": m, "sd": sd, "ci95": half, "runs": int(n), "min": float(a.min())}


def ratio_ci(num, den, seed: int = 0, draws: int = 2000) -> dict:
    """mean(num) / mean(den) with a 95% bootstrap interval (runs resampled with replacement)."""
    a, b = np.asarray(num, np.float64), np.asarray(den, np.float64)
    rng = np.random.default_rng(seed)
    ra = a[rng.integers(0, a.size, (draws, a.size))].mean(axis=1)
    rb = b[rng.integers(0, b.size, (draws, b.size))].mean(axis=1)
    r = ra / rb
    return {"ratio": float(a.mean() / b.mean()), "lo": float(np.percentile(r, 2.5)), "hi": float(np.percentile(r, 97.5))}


def roofline_entry(bytes_per_launch: float, avg_us: float, traffic=None, note: str = "") -> dict:
    a = bytes_per_launch / (avg_us * 1e-6) / 1e9
    out = {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
           "traffic": traffic, "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_us,
           "traffic_frac": (traffic / (avg_us * 1e-6) / 1e9 / HBM_PEAK_GBS) if traffic else None}
    if note:
        out["note"] = note
    return out


def tc_top_bits() -> int:
    """The width of tc_bitmap16's masks of the top ids, as the library reads it (GC_TC_TOP,
    default 4096; 0: no masks)."""
    env = os.environ.get("GC_TC_TOP")
    return int(env) if env is not None else 4096


def tc_env(name: str, default: int) -> int:
    env = os.environ.get(name)
    return int(env) if env is not None else default


def tc_stored_width_bytes(rl: dict, parts: int = 4, split_min: int | None = None, hot_bits: int = 32768,
                          top_bits: int | None = None, mid: bool | None = None, cost: int | None = None,
                          unit: int | None = None) -> dict:
    """Bytes the default triangle count reads per launch (tc_bitmap16, then tc_mid16), by the model
    that fits the kernels (on the degree-ordered DAG they run on, `rl` = GpuGraph.relabelled_numpy()).

    Every DAG edge u -> v is counted on one side. tc_bitmap16's side: L(v) at its stored width (its
    entries among the top `hot_bits` ids as 16-bit offsets in whole 16-byte octs of eight, its other
    entries at 4 bytes; a list whose entries among the top `top_bits` ids fill more octs than their
    mask holds the mask instead, from the mask's oct of id v + 1 on), plus, once per launch, the
    offsets and the DAG split, 4(n+1) + 4n, and L(u) at 4 bytes per entry once per part it is split
    into (rows of at least `split_min` DAG entries run in `parts` parts, shorter rows once) for the
    rows with an edge left on this side, and one bit per DAG entry (which side). tc_mid16's side
    (`mid`, the library's default; an edge u -> v with v among the hot ids whose walk is empty or
    whose walk's octs plus `cost` are fewer than the octs of L(v)): the walk, the 16-byte octs of
    the part of L(u) above v in natural order; an 8-byte record per entry of the hot rows; and L(v)
    at 4 bytes twice (its bits set, then cleared) per work unit of `unit` entries of v's row."""
    top_bits = tc_top_bits() if top_bits is None else top_bits
    mid = (os.environ.get("GC_TC_MID", "1") != "0") if mid is None else mid
    cost = tc_env("GC_TC_MID_COST", 2) if cost is None else cost
    unit = max(64, tc_env("GC_TC_MID_UNIT", 8192)) if unit is None else unit
    split_min = tc_env("GC_TC_SPLIT", 1024 if mid else 256) if split_min is None else split_min
    off = rl["offsets"].astype(np.int64)
    fwd = rl["fwd"].astype(np.int64)
    tgt = rl["targets"]
    n = off.size - 1
    hot_base = ((n - hot_bits) + 31) & ~31 if n > hot_bits else 0
    dplus = off[1:] - fwd
    # DAG entries: targets[fwd[u] .. off[u+1]) of every row u
    row = np.repeat(np.arange(n, dtype=np.int32), dplus)
    dag_first = np.concatenate(([0], np.cumsum(dplus)[:-1]))
    starts = np.repeat(fwd - dag_first, dplus)
    dag_v = tgt[np.arange(row.size, dtype=np.int64) + starts].astype(np.int64)
    del starts
    hot = np.bincount(row, weights=(dag_v >= hot_base), minlength=n).astype(np.int64)
    cold = dplus - hot
    octs = (hot + 7) // 8
    masked_rows = 0
    if top_bits:
        top_octs = top_bits // 128
        top_base = max(hot_base, ((n - top_bits) + 31) & ~31 if n > top_bits else 0)
        top = np.bincount(row, weights=(dag_v >= top_base), minlength=n).astype(np.int64)
        # a mask starts at the oct of id max(v + 1, top_base): L(v) holds ids above v only
        ids = np.arange(n, dtype=np.int64)
        mask_octs = top_octs - np.where(ids + 1 > top_base, (ids + 1 - top_base) >> 7, 0)
        masked = (top >> 3) > mask_octs
        octs = np.where(masked, (hot - top + 7) // 8 + mask_octs, octs)
        masked_rows = int(masked.sum())
    mid_bytes = 0
    mid_edges = 0
    if mid:
        k = np.arange(row.size, dtype=np.int64) - np.repeat(dag_first, dplus)  # place in L(u)
        last = k == dplus[row] - 1
        # octs from the one holding the entry after v to the end of u's hot part
        walk = np.where(last, 0, (hot[row] + 7) // 8 - (k - cold[row] + 1) // 8)
        is_mid = (dag_v >= hot_base) & (last | (walk + cost < octs[dag_v]))
        walk_octs = int(walk[is_mid].sum())
        del walk
        units = np.zeros(n, np.int64)
        deg = np.diff(off)
        units[hot_base:] = (deg[hot_base:] + unit - 1) // unit
        mid_bytes = 16 * walk_octs + 8 * int(off[n] - off[hot_base])
        if os.environ.get("GC_TC_MID32", "0") != "0":
            # the 32-bit tier (tc_mid32): v in [mid_base, hot_base), walks of 32-bit entries
            ids = max(1, tc_env("GC_TC_MID32_IDS", 131072))
            mid_base = n - ids if n > ids else 0
            suffix = dplus[row] - k - 1
            band = (dag_v >= mid_base) & (dag_v < hot_base)
            is32 = band & (last | ((suffix + 3) // 4 + cost < octs[dag_v] + (cold[dag_v] + 3) // 4))
            mid_bytes += 4 * int(suffix[is32].sum()) + 8 * int(off[hot_base] - off[mid_base])
            units[mid_base:hot_base] = (deg[mid_base:hot_base] + unit - 1) // unit
            is_mid |= is32
            del suffix, band, is32
        del k, last
        mid_edges = int(is_mid.sum())
        mid_bytes += int((8 * dplus * units).sum())
        row_side = ~is_mid
        del is_mid
    else:
        row_side = np.ones(row.size, bool)
    indeg = np.bincount(dag_v[row_side], minlength=n).astype(np.int64)
    streamed = int((indeg * (16 * octs + 4 * cold)).sum())
    rowwork = (dplus >= 2) & (np.bincount(row[row_side], minlength=n) > 0)
    row_reads = int((4 * dplus * np.where(dplus >= split_min, parts, 1) * rowwork).sum())
    row_total = 4 * (n + 1) + 4 * n + row_reads + streamed + (row.size // 8 if mid else 0)
    indeg_all = np.bincount(dag_v, minlength=n).astype(np.int64)
    return {"bytes": row_total + mid_bytes, "row_side_bytes": row_total, "mid_side_bytes": mid_bytes,
            "mid": bool(mid), "mid_edges": mid_edges, "dag_edges": int(row.size), "list_bytes": streamed,
            "list_bytes_32bit": int((indeg * 4 * dplus).sum()), "row_bytes": row_reads, "hot_base": int(hot_base),
            "top_bits": int(top_bits), "masked_rows": masked_rows, "split_min": int(split_min),
            "hot_fraction_of_streamed_entries": float((indeg_all * hot).sum() / max(1, (indeg_all * dplus).sum()))}


def ktruss_support_bytes(rl: dict, parts: int = 4, split_min: int = 256, hot_bits: int = 32768) -> dict:
    """Bytes of one k-truss support pass by the model that fits its kernels (on the degree-ordered
    DAG, `rl` = GpuGraph.relabelled_numpy()): kt_bitmap_support reads the offsets and the DAG split,
    4(n+1) + 4n, L(u) at 4 bytes once per part for the rows with an edge on its side, one bit per
    DAG entry, and L(v) at 4 bytes for every DAG edge u->v on its side; kt_mid_support (unless
    GC_KTRUSS_MID=0) takes the edges u->v with v among the hot ids whose walk of L(u) above v, plus
    GC_KTRUSS_MID_COST entries (16), is shorter than L(v) or empty, and reads their walks at 4 bytes
    an entry, an 8-byte record per entry of the hot rows, and L(v) twice per work unit."""
    mid = os.environ.get("GC_KTRUSS_MID", "1") != "0"
    cost = tc_env("GC_KTRUSS_MID_COST", 16)
    unit = max(64, tc_env("GC_KTRUSS_MID_UNIT", 8192))
    off = rl["offsets"].astype(np.int64)
    fwd = rl["fwd"].astype(np.int64)
    tgt = rl["targets"]
    n = off.size - 1
    hot_base = ((n - hot_bits) + 31) & ~31 if n > hot_bits else 0
    dplus = off[1:] - fwd
    row = np.repeat(np.arange(n, dtype=np.int32), dplus)
    dag_first = np.concatenate(([0], np.cumsum(dplus)[:-1]))
    dag_v = tgt[np.arange(row.size, dtype=np.int64) + np.repeat(fwd - dag_first, dplus)].astype(np.int64)
    mid_bytes = 0
    if mid and hot_base < n:
        k = np.arange(row.size, dtype=np.int64) - np.repeat(dag_first, dplus)
        suffix = dplus[row] - k - 1
        del k
        is_mid = (dag_v >= hot_base) & ((suffix == 0) | (suffix + cost < dplus[dag_v]))
        units = np.zeros(n, np.int64)
        deg = np.diff(off)
        units[hot_base:] = (deg[hot_base:] + unit - 1) // unit
        mid_bytes = (4 * int(suffix[is_mid].sum()) + 8 * int(off[n] - off[hot_base]) + int((8 * dplus * units).sum()))
        row_side = ~is_mid
        del suffix, is_mid
    else:
        row_side = np.ones(row.size, bool)
    streamed = int((4 * dplus[dag_v[row_side]]).sum())
    rowwork = np.bincount(row[row_side], minlength=n) > 0
    row_reads = int((4 * dplus * np.where(dplus >= split_min, parts, 1) * rowwork).sum())
    row_bytes = 4 * (n + 1) + 4 * n + row_reads + streamed + (row.size // 8 if mid else 0)
    return {"bytes": row_bytes + mid_bytes, "row_side_bytes": row_bytes, "mid_side_bytes": mid_bytes}


def dag_merge_bytes(off: np.ndarray, tgt: np.ndarray) -> int:
    """The merge model of one symmetric graph's triangle count: 4(n+1) + 4 m_o + 4 sum over DAG
    edges u->v of (d+(u) + d+(v)), on the degree-ordered DAG (u -> v when v ranks above u by
    (degree, id))."""
    off = off.astype(np.int64)
    n = off.size - 1
    deg = np.diff(off)
    order = np.argsort(deg, kind="stable")
    rank = np.empty(n, np.int64)
    rank[order] = np.arange(n)
    row = np.repeat(np.arange(n), deg)
    up = rank[tgt.astype(np.int64)] > rank[row]
    dplus = np.bincount(row[up], minlength=n)
    u, v = row[up], tgt[up].astype(np.int64)
    return int(4 * (n + 1) + 4 * up.sum() + 4 * (dplus[u] + dplus[v]).sum())


def sum_kernels(dev, family: str) -> float:
    return sum(dev.kernel_time(k)[0] for k in KERNEL_FAMILIES[family])


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
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on",
                    help="the C oracle on the host (bounded sample) and the PageRank precision study")
    ap.add_argument("--kernel-events", choices=["on", "all", "off"], default="on",
                    help="time the roofline kernels (on) or every kernel launch (all) of the timed "
                         "steps with HIP events")
    ap.add_argument("--batched", choices=["on", "off"], default="on",
                    help="also time the batched schedules against the one-at-a-time ones: 128-source "
                         "multi-source BFS vs 128 single-source BFS, and the 256-graph triangle-count "
                         "batch (default and autotuned shape) vs one call per graph")
    ap.add_argument("--scheduler", choices=["on", "off", "quick"], default="on",
                    help="the learned schedules against the fixed rules and the forced choices on R-MAT "
                         "22 (this graph), R-MAT 24 and the 4096 x 4096 grid (quick: this graph only)")
    ap.add_argument("--reference-convention", choices=["on", "off"], default="on",
                    help="the reference's timing convention: host outputs, best of 5 after 1 warmup")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: the host threads this process may use (OMP_NUM_THREADS, else all)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes 5 times (the known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings) and exit")
    ap.add_argument("--workload", choices=["step", "scheduler"], default="step",
                    help="step: the R-MAT step above; scheduler: only the learned-versus-fixed "
                         "comparison on R-MAT 24 and the 4096 x 4096 grid")
    ap.add_argument("--model", default=None, help="learned schedules (default: the committed model)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
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

    if args.workload == "scheduler":
        dev = grafeo_cuda.GpuDevice(0)
        out = {"metric": "learned vs fixed schedule (BFS per level, TC per degree bucket)", "unit": "ms",
               "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
               "scheduler": scheduler_section(dev, None, None, args, with_large=True)}
        print(json.dumps(out), flush=True)
        return 0

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

    # ---- rooflines (HBM bandwidth for every kernel here; MFMA is not applicable) ----
    # PageRank SpMV per iteration: offsets + targets + x read once + y written once.
    spmv_bytes = 4 * (n + 1) + 4 * m + 4 * n + 4 * n
    # BFS per traversal: offsets + targets read once, distances written once.
    bfs_bytes = 4 * (n + 1) + 4 * m + 4 * n
    # The merge model of triangle counting (tc_hash, tc_count): offsets, the DAG's entries, and
    # for each DAG edge u->v both out-lists at 4 bytes an entry.
    tc_merge_bytes = 4 * (n + 1) + 4 * m_o.value + 4 * inter.value
    wedge_bytes = 4 * (n + 1) + 4 * n + 4 * m + 4 * wedges.value
    t0 = time.time()
    tc16 = tc_stored_width_bytes(g.relabelled_numpy(), mid=os.environ.get("GC_TC_MID", "1") != "0"
                                 and os.environ.get("GC_TC_HOT16", "1") != "0")
    log(f"tc_bitmap16 model: {tc16['bytes'] / 1e9:.2f} GB per launch ({time.time() - t0:.1f}s)")
    traffic = load_traffic(f"rmat{args.scale}_ef{args.edge_factor}_seed{args.seed}")
    rooflines = {}
    # PageRank launches in batches and the launches after convergence return at once: its
    # SpMV time per iteration is the kernel's time per step over the iterations that ran.
    for name in ("pr_spmv_merge", "pr_spmv_push", "pr_spmv_hot", "pr_spmv"):
        if name in kernels:
            per_iter_us = 1000.0 * kernels[name]["total_ms"] / max(1, iters)
            t = traffic.get(name)
            t = t * kernels[name]["launches"] / max(1, iters) if t else None
            rooflines[name] = roofline_entry(
                spmv_bytes, per_iter_us, t,
                "bytes_per_launch = 4(n+1) + 4m + 8n (offsets, targets, x read once, y written once); "
                "avg_launch_us is per iteration that ran (launches after convergence return at once)"
                + ("; the push-pull kernels read the hubs' lists only inside the push set (79% of the "
                   "target stream at R-MAT 22)" if name in ("pr_spmv_merge", "pr_spmv_push") else ""))
    bin_set = next((ks for ks in PR_BIN_SETS if all(k in kernels for k in ks)), None)
    if bin_set is not None:
        lay = pagerank_layout(g)
        zero_rows = int(np.count_nonzero(np.diff(offsets) == 0))
        models = pagerank_bin_models(lay, zero_rows)
        per_iter = {k: 1000.0 * kernels[k]["total_ms"] / max(1, iters) for k in bin_set}
        tr = [traffic.get(k) for k in bin_set]
        t_sum = (sum(t * kernels[k]["launches"] / max(1, iters) for t, k in zip(tr, bin_set))
                 if all(tr) else None)
        rooflines["pr_spmv_binned"] = roofline_entry(
            spmv_bytes, sum(per_iter.values()), t_sum,
            "the SpMV of the default PageRank (binned cold entries): " + " + ".join(bin_set)
            + " per iteration that ran, against bytes_per_launch = 4(n+1) + 4m + 8n (offsets, "
            "targets, x read once, y written once)")
        rooflines["pr_spmv_binned"]["kernels_us"] = per_iter
        rooflines["pr_spmv_binned"]["layout"] = lay
        for k in bin_set:
            t = traffic.get(k)
            rooflines[k] = roofline_entry(
                models[k], per_iter[k], t * kernels[k]["launches"] / max(1, iters) if t else None,
                "stored-width model of the binned layout (bench.py pagerank_bin_models); avg_launch_us is "
                "per iteration that ran")
    for name in ("tc_hash", "tc_count"):
        if name in kernels:
            rooflines[name] = roofline_entry(tc_merge_bytes, kernels[name]["avg_us"], traffic.get(name),
                                             "merge model: 4(n+1) + 4 m_o + 4 sum over DAG edges (d+u + d+v)")
    tc_note = ("stored-width model (bench.py tc_stored_width_bytes): 4(n+1) + 4n, L(u) at 4 bytes once per "
               "part it is split into, and for every DAG edge u->v counted by tc_bitmap16 the list L(v) at its "
               "stored width (entries among the top 32768 ids as 16-bit offsets in 16-byte octs of eight, "
               f"others at 4 bytes, the entries among the top {tc16['top_bits']} ids as a mask of at most "
               f"{tc16['top_bits'] // 8} bytes where that is smaller: {tc16['masked_rows']} lists)")
    if tc16["mid"]:
        tc_note += (f"; the {tc16['mid_edges']} of {tc16['dag_edges']} DAG edges u->v counted by tc_mid16 at v: "
                    "the octs of L(u) above v (16-byte octs, natural order; 4-byte entries in the 32-bit tier "
                    "when it is on), an 8-byte record per entry of the rows it serves, L(v) at 4 bytes twice "
                    "per work unit, and one bit per DAG entry")
    tc_note += f"; the merge model (both lists at 4 bytes, {tc_merge_bytes} bytes) would count data never read"
    for name in ("tc_bitmap16", "tc_bitmap"):
        if name in kernels:
            # tc_bitmap streams the same lists at 4 bytes per entry
            model = tc16["row_side_bytes"] if name == "tc_bitmap16" else (
                tc16["bytes"] - tc16["list_bytes"] + tc16["list_bytes_32bit"])
            rooflines[name] = roofline_entry(model, kernels[name]["avg_us"], traffic.get(name), tc_note)
            rooflines[name]["merge_model_bytes"] = tc_merge_bytes
    if "tc_mid16" in kernels and "tc_bitmap16" in kernels:
        mids = [k for k in ("tc_mid16", "tc_mid32") if k in kernels]
        mid_us = sum(kernels[k]["avg_us"] for k in mids)
        rooflines["+".join(mids)] = roofline_entry(tc16["mid_side_bytes"], mid_us, None, tc_note)
        # the count as one unit: every kernel of it, once each per count
        pair = ["tc_bitmap16"] + mids
        t_pair = [traffic.get(k) for k in pair]
        rooflines["+".join(pair)] = roofline_entry(
            tc16["bytes"], sum(kernels[k]["avg_us"] for k in pair), sum(t_pair) if all(t_pair) else None, tc_note)
        rooflines["+".join(pair)]["merge_model_bytes"] = tc_merge_bytes
    if "tc_wedge" in kernels:
        rooflines["tc_wedge"] = roofline_entry(wedge_bytes, kernels["tc_wedge"]["avg_us"], traffic.get("tc_wedge"))
    # BFS over the whole call per source (every level's kernels, launch gaps and the read-back of
    # each level's counters), not kernel time alone.
    rooflines["bfs_traversal"] = roofline_entry(
        bfs_bytes, 1000.0 * per_algo["bfs_ms_per_source"], bfs_call_traffic(
            f"rmat{args.scale}_ef{args.edge_factor}_seed{args.seed}"),
        "4(n+1) + 4m + 4n per traversal; avg_launch_us is the whole BFS call per source")
    # The dominant kernel: the one with the most event time in the profiled step (the triangle
    # count's two kernels, which split its edges between them, as one).
    units = {k: v["total_ms"] for k, v in kernels.items()}
    if "tc_bitmap16" in units and "tc_mid16" in units:
        pair = ["tc_bitmap16"] + [k for k in ("tc_mid16", "tc_mid32") if k in units]
        units["+".join(pair)] = sum(units.pop(k) for k in pair)
    dominant = max(units, key=units.get) if units else None
    roofline = dict(rooflines.get(dominant) or {})
    if roofline:
        roofline["kernel"] = dominant

    parity = {}
    try:
        with open(GOLDEN) as f:
            golden = json.load(f)
        key = f"rmat{args.scale}_ef{args.edge_factor}_seed{args.seed}"
        if key in golden:
            parity["triangles_match_golden"] = results[-1]["triangles"] == golden[key]["triangles"]
    except FileNotFoundError:
        pass
    bfs_golden = os.path.join(REPO, "tests", "golden", "rmat22_bfs.json")
    if args.scale == 22 and args.edge_factor == 16 and args.seed == 1 and os.path.exists(bfs_golden):
        with open(bfs_golden) as f:
            want = json.load(f)["sources"]
        host = np.empty(n, np.int32)
        ok = True
        for s in sources:
            check(lib.gc_bfs_device(g._handle, s, d_dist, byref(c_uint32()), byref(c_int32()), byref(c_double())))
            check(lib.gc_device_copy_to_host(dev._handle, host.ctypes.data_as(c_void_p), d_dist, 4 * n))
            ok = ok and want.get(str(s), {}).get("sha256") == hashlib.sha256(host.tobytes()).hexdigest()
        parity["bfs_distances_match_golden"] = ok

    reference = None
    if args.reference_convention == "on":
        reference = reference_convention(dev, g, offsets, sources, args, m)
        rooflines.update(reference.pop("rooflines"))

    batched = None
    if args.batched == "on":
        batched = batched_section(dev, g, offsets, n, m, args)
        rooflines.update(batched.pop("rooflines"))

    sched = None
    if args.scheduler != "off":
        sched = scheduler_section(dev, g, offsets, args, with_large=args.scheduler == "on")

    cpu, precision = None, None
    if args.cpu_baseline == "on":
        cpu = cpu_baseline(offsets, targets, sources, iters, args.cpu_threads, m)
        precision = pagerank_precision(g, offsets, targets, dev)

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
        "reference_convention": reference,
        "cpu_baseline": cpu,
        "pagerank_precision": precision,
        "batched": batched,
        "scheduler": sched,
        "device": dev.name,
    }
    check(lib.gc_device_free(dev._handle, d_dist))
    check(lib.gc_device_free(dev._handle, d_rank))
    print(json.dumps(out), flush=True)
    return 0


def reference_convention(dev, g, offsets, sources, args, m) -> dict:
    """The reference's timing convention (crates/grafeo-cuda-algos/src/bin/gpu_*.rs): each
    algorithm's public call with host outputs (the distances and ranks copied back inside the
    call, as its time_seconds includes them, bfs.rs:121-130, pagerank.rs:190), warmup 1, runs 5,
    the best run; BFS TEPS per source and their harmonic mean over the sources
    (gpu_bfs.rs:116-124), plus a Graph500-style figure (undirected edges of the reached
    component over the time). Also the k-truss (k = 4) call and its roofline."""
    deg = np.diff(offsets.astype(np.int64))

    def best(call, warmup=1, runs=5):
        for _ in range(warmup):
            call()
        return min(call() for _ in range(runs))

    out = {"convention": "host outputs, best of 5 calls after 1 warmup"}
    out["tc"] = {"time_s": best(lambda: g.triangle_count()["time_seconds"])}
    out["tc"]["gteps"] = m / out["tc"]["time_s"] / 1e9
    per_source, g500 = [], []
    for s in sources:
        last = {}

        def call():
            r = g.bfs_numpy(s)
            last["r"] = r
            return r["time_seconds"]

        t = best(call)
        per_source.append(t)
        reached = last["r"]["distances"] >= 0
        g500.append(float(deg[reached].sum()) / 2.0 / t)
    out["bfs"] = {"time_s_per_source": per_source,
                  "gteps_harmonic_mean": len(sources) / sum(t / m for t in per_source) / 1e9,
                  "graph500_gteps_harmonic_mean": len(g500) / sum(1.0 / x for x in g500) / 1e9}
    pr = {}

    def pr_call():
        r = g.pagerank_numpy(0.85, args.pr_iterations, 1e-6)
        pr["it"] = r["iterations"]
        return r["time_seconds"]

    out["pagerank"] = {"time_s": best(pr_call)}
    out["pagerank"]["iterations"] = pr["it"]
    out["pagerank"]["gteps"] = m * pr["it"] / out["pagerank"]["time_s"] / 1e9
    # k-truss, k = 4 (config of the R-MAT 20 suite; gpu_ktruss.rs:51-56 reports entries / time)
    kt = {}

    def kt_call():
        r = g.k_truss(4)
        kt["r"] = r
        return r["time_seconds"]

    dev.profile(True, only=None)
    out["ktruss"] = {"k": 4, "time_s": best(kt_call, warmup=1, runs=5)}
    kt_kernel_ms = sum_kernels(dev, "ktruss") / 6.0
    dev.profile(False)
    rounds = kt["r"]["iterations"]
    out["ktruss"].update({"gteps": m / out["ktruss"]["time_s"] / 1e9, "rounds": rounds,
                          "edges": len(kt["r"]["edges"]), "kernel_ms_per_call": kt_kernel_ms})
    # The survey's model: triangle-count bytes per support pass times passes, plus 5m per peel
    # round. This library counts the supports once, by the triangle walk on the hub bitmap over the
    # 32-bit lists (tc_bitmap's stored bytes, plus a 4-byte read-modify-write of one support per
    # triangle), and then only peels.
    ktm = ktruss_support_bytes(g.relabelled_numpy(), split_min=tc_env("GC_TC_SPLIT", 256))
    triangles = g.triangle_count()["total_triangles"]
    kt_bytes = ktm["bytes"] + 8 * triangles + 5 * m * rounds
    out["rooflines"] = {"ktruss": roofline_entry(
        kt_bytes, 1000.0 * kt_kernel_ms,
        note="one support pass (bench.py ktruss_support_bytes: each DAG edge's supports walked from its "
             f"cheaper end, {ktm['row_side_bytes']} bytes on the row side, {ktm['mid_side_bytes']} on the middle "
             "side) + 8 bytes per triangle, plus 5m per peel round; avg_launch_us is the kernels' event time "
             "per k-truss call (k = 4)")}
    log(f"reference convention: {json.dumps({k: v for k, v in out.items() if k != 'rooflines'})[:400]}")
    return out


def batched_section(dev, g, offsets, n, m, args) -> dict:
    """Batched schedules against the same work done one unit at a time, on the same GPU, same
    inputs, bit-identical results (checked here); speed-ups by call time and by kernel time."""
    import grafeo_cuda
    from grafeo_cuda import workloads
    from grafeo_cuda._native import check, lib

    out = {"rooflines": {}}
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
    dev.profile(True, only=None)
    run_ms()
    ms_kernel_ms = sum_kernels(dev, "msbfs") + sum_kernels(dev, "util")
    dev.profile(False)
    run_single()
    single_t, reach1 = min((run_single() for _ in range(2)), key=lambda x: x[0])
    dev.profile(True, only=None)
    run_single()
    single_kernel_ms = sum_kernels(dev, "bfs") + sum_kernels(dev, "util")
    dev.profile(False)
    # row 0 of the batched distances against the single-source distances of source 0
    row = np.empty(n, np.int32)
    one = np.empty(n, np.int32)
    check(lib.gc_device_copy_to_host(dev._handle, row.ctypes.data_as(c_void_p), d_ms, 4 * n))
    rr, dd = c_uint32(), c_int32()
    check(lib.gc_bfs_device(g._handle, sources[0], d_one, byref(rr), byref(dd), byref(secs)))
    check(lib.gc_device_copy_to_host(dev._handle, one.ctypes.data_as(c_void_p), d_one, 4 * n))
    check(lib.gc_device_free(dev._handle, d_ms))
    check(lib.gc_device_free(dev._handle, d_one))
    levels = int(depth.max()) + 1
    out["msbfs"] = {
        "sources": len(sources), "ms": 1000 * ms_t, "gteps": m * len(sources) / ms_t / 1e9,
        "kernel_ms": ms_kernel_ms,
        "single_source_ms": 1000 * single_t, "single_source_gteps": m * len(sources) / single_t / 1e9,
        "single_source_kernel_ms": single_kernel_ms,
        "speedup": single_t / ms_t, "speedup_kernel_time": single_kernel_ms / ms_kernel_ms,
        "levels": levels,
        "identical": bool(list(reach) == reach1 and np.array_equal(row, one)),
    }
    # Model that fits the multi-source kernels: per level the offsets, the adjacency once and the
    # three 16-byte masks per vertex (seen, visit, next), plus the distances of every source once.
    ms_bytes = levels * (4 * (n + 1) + 4 * m + 48 * n) + 4 * len(sources) * n
    out["rooflines"]["msbfs"] = roofline_entry(
        ms_bytes, 1000.0 * ms_kernel_ms,
        note="levels x (4(n+1) + 4m + 48n) + 4 S n (structure and three 16-byte masks per level, "
             "distances once); avg_launch_us is the kernels' event time per call; the per-source sum "
             f"of the BFS model would be {len(sources) * (4 * (n + 1) + 8 * n + 4 * m)} bytes")
    log(f"msbfs: {out['msbfs']}")

    # -- 256-graph triangle-count batch vs one call per graph --
    t0 = time.time()
    graphs = workloads.uneven_batch(256, 1)
    entries = sum(int(t.size) for _, t in graphs)
    batch = grafeo_cuda.GpuBatch.from_csrs(graphs, dev)
    singles = [grafeo_cuda.GpuGraph.from_csr(o, t, dev) for o, t in graphs]
    for s_g in singles:
        s_g.prepare()
    model_bytes = sum(dag_merge_bytes(o, t) for o, t in graphs)
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
    dev.profile(True, only=["bt_count"])
    for _ in range(5):
        batch.triangle_count()
    bt_ms, bt_n = dev.kernel_time("bt_count")
    dev.profile(False)
    bt_kernel_ms = bt_ms / max(1, bt_n)

    def time_singles():
        total, counts = 0.0, []
        for s_g in singles:
            r = s_g.triangle_count()
            total += r["time_seconds"]
            counts.append(r["total_triangles"])
        return total, counts

    time_singles()
    single_t, counts_single = min((time_singles() for _ in range(2)), key=lambda x: x[0])
    dev.profile(True, only=None)
    time_singles()
    singles_kernel_ms = sum_kernels(dev, "tc") + sum_kernels(dev, "util")
    dev.profile(False)
    out["batch_tc"] = {
        "graphs": len(graphs), "csr_entries": entries,
        "default_config": list(default_cfg), "default_ms": 1000 * default_t,
        "tuned_config": list(tuned["config"]), "tuned_ms": 1000 * tuned_t,
        "tuned_gteps": entries / tuned_t / 1e9, "kernel_ms": bt_kernel_ms,
        "one_call_per_graph_ms": 1000 * single_t, "one_call_per_graph_kernel_ms": singles_kernel_ms,
        "speedup_vs_one_call_per_graph": single_t / tuned_t,
        "speedup_kernel_time": singles_kernel_ms / bt_kernel_ms,
        "speedup_tuned_vs_default": default_t / tuned_t,
        "identical": counts_default == counts_tuned == counts_single,
        "match_golden": None if expected is None else counts_tuned == expected,
    }
    out["rooflines"]["batch_tc"] = roofline_entry(
        model_bytes, 1000.0 * bt_kernel_ms,
        note="the sum over the 256 graphs of the merge model 4(n+1) + 4 m_o + 4 sum (d+u + d+v); "
             "avg_launch_us is bt_count's event time per replay")
    log(f"batch_tc: {out['batch_tc']}")
    return out


def forced_scheduler(dev, choice: int):
    """A network whose outputs make `choice` the cheapest wherever it is allowed."""
    import grafeo_cuda

    s = grafeo_cuda.GpuScheduler(dev, seed=1)
    p = np.zeros(1571, np.float32)
    # ln-microsecond outputs 10 apart: far above the triangle count's 20 us switching margin
    b3 = np.full(3, 10.0, np.float32)
    b3[choice] = 0.0
    p[-3:] = b3
    s.set(params=p)
    return s


def scheduler_section(dev, g22, off22, args, with_large: bool) -> dict:
    """The learned schedules against the fixed rules on the same GPU and inputs: per graph, the
    BFS total time per source (mean and 95% confidence interval over 5 runs after a warmup, each
    run over all the graph's sources), the time per level, and the triangle count under each
    schedule; distances and counts must be identical under every schedule. Rows: Beamer's
    switch alone, Beamer with device-chained top-down levels (the fixed default), each choice
    forced, and the learned network. Graphs: R-MAT 22 (the step's graph, 8 sources), and with
    `with_large` R-MAT 24 (4 sources) and the 4096 x 4096 grid (centre and corner), which the
    network was not trained on."""
    import grafeo_cuda
    from grafeo_cuda import scheduler
    from grafeo_cuda._native import check, lib

    bfs_s, tc_s = scheduler.load(dev, args.model or scheduler.MODEL)
    forced = [forced_scheduler(dev, c) for c in range(3)]
    runs = 5
    out = {"runs": runs, "graphs": {}}

    def graph_cases():
        if g22 is not None:
            yield "rmat22", g22, off22, pick_sources(off22, args.sources, args.seed)
        if with_large:
            for name, pairs_fn, n in (("rmat24", lambda: grafeo_cuda._native.gen_rmat(24, 16, 1), 1 << 24),
                                      ("grid4096", lambda: grafeo_cuda._native.gen_grid(4096, 4096), 4096 * 4096)):
                t0 = time.time()
                pairs = pairs_fn()
                gg = grafeo_cuda.GpuGraph.from_edge_pairs(pairs, n, dev)
                del pairs
                gg.prepare()
                off = np.empty(n + 1, np.uint32)
                check(lib.gc_device_copy_to_host(dev._handle, off.ctypes.data_as(c_void_p),
                                                 c_void_p(lib.gc_graph_device_offsets(gg._handle)), 4 * (n + 1)))
                srcs = pick_sources(off, 4, args.seed) if name == "rmat24" else [2048 * 4096 + 2048, 0]
                log(f"{name}: n={n} m={gg.num_edges} ready in {time.time() - t0:.1f}s")
                yield name, gg, off, srcs
                del gg

    bfs_rows = [("beamer", {"GC_BFS_RULE": "beamer"}, None), ("beamer+chain (fixed default)", {}, None),
                ("forced top-down", {}, forced[0]), ("forced bottom-up", {}, forced[1]),
                ("forced chained", {}, forced[2]), ("learned", {}, bfs_s)]
    tc_rows = [("bitmap (fixed)", None), ("forced hash", forced[1]), ("forced lists", forced[2]),
               ("learned", tc_s)]
    for name, g, off, srcs in graph_cases():
        n = off.size - 1
        d_dist = c_void_p()
        check(lib.gc_device_malloc(dev._handle, 4 * n, byref(d_dist)))
        host = np.empty(n, np.int32)
        entry = {"num_nodes": n, "csr_entries": g.num_edges, "bfs_sources": len(srcs), "bfs": {}, "tc": {}}
        reference_hashes = None
        for label, env, sched in bfs_rows:
            os.environ.pop("GC_BFS_RULE", None)
            os.environ.update(env)
            g.set_scheduler(bfs=sched)
            hashes, levels = [], 0
            for s in srcs:  # warmup, with the distances checked
                reach, depth, secs = c_uint32(), c_int32(), c_double()
                check(lib.gc_bfs_device(g._handle, s, d_dist, byref(reach), byref(depth), byref(secs)))
                check(lib.gc_device_copy_to_host(dev._handle, host.ctypes.data_as(c_void_p), d_dist, 4 * n))
                hashes.append((reach.value, depth.value, hashlib.sha256(host.tobytes()).hexdigest()))
                levels += depth.value + 1
            if reference_hashes is None:
                reference_hashes = hashes
            samples = []
            for _ in range(runs):
                total = 0.0
                for s in srcs:
                    secs = c_double()
                    check(lib.gc_bfs_device(g._handle, s, d_dist, byref(c_uint32()), byref(c_int32()), byref(secs)))
                    total += secs.value
                samples.append(1000.0 * total / len(srcs))
            st = mean_ci(samples)
            entry["bfs"][label] = {"ms_per_source": st, "levels_per_source": levels / len(srcs),
                                   "us_per_level": 1000.0 * st["mean"] * len(srcs) / levels,
                                   "samples_ms": samples, "identical": hashes == reference_hashes,
                                   "last_schedule": {str(c): g.last_schedule().count(c) for c in (0, 1, 2)}}
            log(f"{name} BFS {label}: {st['mean']:.3f} +- {st['ci95']:.3f} ms/source")
        os.environ.pop("GC_BFS_RULE", None)
        g.set_scheduler()
        b = entry["bfs"]
        entry["bfs_learned_speedup"] = {
            other: ratio_ci(b[other]["samples_ms"], b["learned"]["samples_ms"])
            for other in ("beamer", "beamer+chain (fixed default)", "forced top-down", "forced bottom-up",
                          "forced chained")}
        counts = None
        for label, sched in tc_rows:
            g.set_scheduler(tc=sched)
            first = g.triangle_count()["total_triangles"]
            samples = [1000.0 * g.triangle_count()["time_seconds"] for _ in range(runs)]
            counts = first if counts is None else counts
            entry["tc"][label] = {"ms": mean_ci(samples), "samples_ms": samples, "identical": first == counts,
                                  "triangles": first, "schedule": g.last_schedule()}
            log(f"{name} TC {label}: {entry['tc'][label]['ms']['mean']:.2f} ms")
        g.set_scheduler()
        entry["tc_learned_speedup_vs_fixed"] = ratio_ci(entry["tc"]["bitmap (fixed)"]["samples_ms"],
                                                        entry["tc"]["learned"]["samples_ms"])
        check(lib.gc_device_free(dev._handle, d_dist))
        out["graphs"][name] = entry
    return out


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


def host_cpu() -> dict:
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def cpu_baseline(offsets, targets, sources, iters, threads, m) -> dict:
    """The C oracle (a port of the reference's semantics, OpenMP) on the host, on a bounded
    sample of the step: the triangle count (its degree-ordered DAG built once, then the merge over
    every stride-th DAG row, times the stride), one BFS source, and PageRank iterations; each
    sample best of 5 after 1 warmup (the reference's defaults, gpu_triangles.rs:14-20), on all the
    threads this process may use and on one thread; scaled to the step's mix."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    threads = threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)

    def best(fn, warmup=1, runs=5):
        for _ in range(warmup):
            fn()
        out = []
        for _ in range(runs):
            t = time.perf_counter()
            fn()
            out.append(time.perf_counter() - t)
        return min(out)

    def measure(th: int, tc_stride: int, pr_its: int) -> dict:
        oracle.set_threads(th)
        t = time.perf_counter()
        sample = oracle.TriangleSample(offsets, targets)
        build = time.perf_counter() - t
        try:
            count = best(lambda: sample.count_rows(tc_stride, 0))
        finally:
            sample.close()
        tc = build + tc_stride * count
        bfs = best(lambda: oracle.bfs(offsets, targets, sources[0]))
        pr = best(lambda: oracle.pagerank(offsets, targets, 0.85, pr_its, 0.0)) / pr_its
        step = tc + len(sources) * bfs + iters * pr
        return {"threads": th, "tc_s": tc, "tc_dag_build_s": build, "tc_row_stride": tc_stride,
                "bfs_s_per_source": bfs, "pagerank_s_per_iteration": pr, "step_s": step,
                "gteps": m * (1 + len(sources) + iters) / step / 1e9}

    t0 = time.time()
    full = measure(threads, 8, 2)
    one = measure(1, 512, 1)
    log(f"cpu baseline ({time.time() - t0:.1f}s): {threads} threads {full['gteps']:.3f} GTEPS, "
        f"1 thread {one['gteps']:.3f} GTEPS")
    return {
        "value": full["gteps"],
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "host": host_cpu(),
        "one_thread": one,
        "all_threads": full,
        "sample": (f"the step's mix from samples, each best of 5 after 1 warmup: triangle count = DAG build "
                   f"+ 8 x the merge over every 8th DAG row ({threads} threads; every 512th row on one "
                   f"thread), 1 BFS source, PageRank 2 iterations (1 on one thread); scaled to 1 triangle "
                   f"count + {len(sources)} BFS + {iters} PageRank iterations"),
    }


def pagerank_precision(g, offsets, targets, dev) -> dict:
    """The PageRank precision study at the step's graph: d = 0.85, at most 50 iterations, tol
    1e-7, FP32 ranks against the FP64 oracle run with the same settings (iteration count and
    convergence must agree), max-abs and L1 error at that iteration, and the SpMV's rate."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    spmv_kernels = ("pr_spmv_merge", "pr_spmv_push", "pr_spmv_hot", "pr_spmv") + PR_BIN_SETS[0] + PR_BIN_SETS[1]
    dev.profile(True, only=list(spmv_kernels))
    r = g.pagerank_numpy(0.85, 50, 1e-7)
    bin_set = next((ks for ks in PR_BIN_SETS if all(dev.kernel_time(k)[1] for k in ks)), None)
    if bin_set is not None:
        name = "+".join(bin_set)
        ms = sum(dev.kernel_time(k)[0] for k in bin_set)
    else:
        name = next(k for k in spmv_kernels if dev.kernel_time(k)[1])
        ms, _ = dev.kernel_time(name)
    dev.profile(False)
    _, it, conv = oracle.pagerank(offsets, targets, 0.85, 50, 1e-7)
    k = r["iterations"]
    ref = oracle.pagerank(offsets, targets, 0.85, k, 0.0)[0]  # the FP64 iterate of the same count
    diff = np.abs(r["ranks"].astype(np.float64) - ref)
    prev = g.pagerank_numpy(0.85, max(0, k - 1), 0.0)["ranks"].astype(np.float64)
    n, m = offsets.size - 1, targets.size
    per_iter_us = 1000.0 * ms / max(1, r["iterations"])
    spmv_bytes = 4 * (n + 1) + 4 * m + 8 * n
    out = {"damping": 0.85, "max_iterations": 50, "tolerance": 1e-7,
           "iterations": k, "converged": r["converged"],
           "oracle_iterations": it, "oracle_converged": conv,
           "iterations_within_one": abs(k - it) <= 1 and bool(r["converged"]) == bool(conv),
           "fp32_l1_change_at_stop": float(np.abs(r["ranks"].astype(np.float64) - prev).sum()),
           "note": "L1 and max-abs against the FP64 oracle's iterate of the same count; the L1 change of "
                   "FP32 iterates has a rounding floor near this tolerance, so the counts may differ by one",
           "l1": float(diff.sum()), "max_abs": float(diff.max()), "l1_gate": 1e-6,
           "spmv_kernel": name, "spmv_us_per_iteration": per_iter_us,
           "spmv_gbs": spmv_bytes / (per_iter_us * 1e-6) / 1e9}
    log(f"pagerank precision: {out}")
    return out


if __name__ == "__main__":
    sys.exit(main())
