"""Training, saving and loading of the learned schedules (an addition with no reference function).

Two networks of the same shape (12 -> 32 -> 32 -> 3, GpuScheduler) make two decisions:
  - BFS, per level: top-down, bottom-up or device-chained levels; trained on
    GpuGraph.bfs_level_timings (every choice timed from the same state at every level);
  - symmetric triangle count, per degree bucket: bitmap, hash or binary-search intersection;
    trained on GpuGraph.tc_bucket_timings.
Targets are ln(time in microseconds); choices that are not allowed are masked out. Inputs are
standardised with the training set's mean and standard deviation. Training runs on the GPU
(gc_sched_train: the library's forward, backward and Adam kernels).

The committed model (sched_model.json next to this file) was produced by

    python3 -m grafeo_cuda.scheduler train --out grafeo-cuda_amd/grafeo_cuda/sched_model.json

on one MI355X, from R-MAT scales 16..22 (edge factor 16) and 4-neighbour grids of 256², 512²,
1024², 2048², 2896², 2048 x 4096 and 2896 x 4096 vertices (the training set below). It is evaluated on graphs it was not trained
on: R-MAT scale 24 and the 4096 x 4096 grid (bench.py --workload scheduler).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import GpuDevice, GpuGraph, GpuScheduler
from . import _native

MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sched_model.json")


def _log(msg: str) -> None:
    print(f"[scheduler] {msg}", file=sys.stderr, flush=True)


def rmat(scale: int, edge_factor: int = 16, seed: int = 1):
    pairs = _native.gen_rmat(scale, edge_factor, seed)
    return _native.from_edge_pairs(pairs, 1 << scale, True)


def grid(rows: int, cols: int):
    return _native.from_edge_pairs(_native.gen_grid(rows, cols), rows * cols, True)


def nonzero_sources(off, count: int, seed: int) -> list[int]:
    deg = np.diff(np.asarray(off, dtype=np.int64))
    cand = np.flatnonzero(deg > 0)
    rng = np.random.default_rng(seed)
    return [int(v) for v in rng.choice(cand, size=min(count, cand.size), replace=False)]


def training_graphs(quick: bool = False):
    """(name, offsets, targets, BFS sources) of the training set."""
    scales = (16, 18) if quick else (16, 17, 18, 19, 20, 21, 22)
    # 2896 x 2896 (8.4M vertices) is the first square size here at which the binary-search
    # intersection beats the bitmap on the grid's degree-4 rows (2571 against 2923 us)
    # and rectangular grids of 8.4 M and 11.9 M vertices, where the bitmap kernels see even
    # fewer of their hot ids
    rows = ((256, 256), (512, 512)) if quick else ((256, 256), (512, 512), (1024, 1024), (2048, 2048),
                                                   (2896, 2896), (2048, 4096), (2896, 4096))
    for s in scales:
        off, tgt = rmat(s, 16, seed=100 + s)
        yield f"rmat{s}", off, tgt, nonzero_sources(off, 3, s)
    for r, c in rows:
        off, tgt = grid(r, c)
        n = r * c
        name = f"grid{r}" if r == c else f"grid{r}x{c}"
        yield name, off, tgt, [0, n // 2 + c // 2, int(np.random.default_rng(n).integers(0, n))]


def _targets(times_us: np.ndarray):
    t = np.asarray(times_us, dtype=np.float64)
    mask = (t > 0).astype(np.float32)
    return np.log(np.maximum(t, 1.0)).astype(np.float32), mask


def collect(device: GpuDevice, graphs) -> dict:
    """Training sets {'bfs': (x, t, w), 'tc': (x, t, w), 'names': [...]} from the kernels' own
    timings. w is the mask of allowed choices scaled so that every graph weighs the same (a
    high-diameter grid gives thousands of BFS levels, an R-MAT graph about ten)."""
    bx, bt, bm, tx, tt, tm, names = [], [], [], [], [], [], []
    for name, off, tgt, sources in graphs:
        t0 = time.time()
        names.append(name)
        g = GpuGraph.from_csr(off, tgt, device)
        g.prepare()
        gx, gt, gm = [], [], []
        for s in sources:
            r = g.bfs_level_timings(s)
            t, m = _targets(r["times_us"])
            gx.append(r["features"])
            gt.append(t)
            gm.append(m)
        levels = sum(len(x) for x in gx)
        bx += gx
        bt += gt
        bm += [m * (100.0 / max(levels, 1)) for m in gm]
        r = g.tc_bucket_timings()
        t, m = _targets(r["times_us"])
        tx.append(r["features"])
        tt.append(t)
        tm.append(m * (100.0 / max(len(t), 1)))
        _log(f"{name}: {sum(len(x) for x in bx)} BFS levels, {sum(len(x) for x in tx)} TC buckets so far "
             f"({time.time() - t0:.1f}s)")
        del g
    return {"bfs": (np.concatenate(bx), np.concatenate(bt), np.concatenate(bm).astype(np.float32)),
            "tc": (np.concatenate(tx), np.concatenate(tt), np.concatenate(tm).astype(np.float32)),
            "names": names}


def centered(t, mask):
    """Targets minus their mean over the allowed choices of each sample. A decision needs only
    the order of its choices' costs; without each state's common level (which grows with the
    graph) the network does not have to extrapolate absolute times to graphs larger than those
    it was trained on."""
    allowed = (mask > 0).astype(np.float64)
    mean = (t * allowed).sum(axis=1) / np.maximum(allowed.sum(axis=1), 1.0)
    return (t - mean[:, None] * allowed).astype(np.float32)


def fit(device: GpuDevice, x, t, mask, steps: int = 3000, lr: float = 3e-3, seed: int = 7):
    """A GpuScheduler trained on (x, t, mask): returns (scheduler, loss before, loss after)."""
    s = GpuScheduler(device, seed=seed)
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    s.set(mean=mean, scale=1.0 / np.where(std > 1e-6, std, 1.0))
    before = s.loss_grad(x, t, mask)[0]
    after = s.train(x, t, mask, steps=steps, lr=lr)
    return s, before, after


def save(path: str, bfs: GpuScheduler, tc: GpuScheduler, info: dict) -> None:
    out = {"format": "grafeo-hip scheduler 1", "info": info}
    for key, s in (("bfs", bfs), ("tc", tc)):
        w = s.get()
        out[key] = {k: [float(v) for v in w[k]] for k in ("params", "mean", "scale")}
    with open(path, "w") as f:
        json.dump(out, f)


def load(device: GpuDevice, path: str = MODEL):
    """(bfs scheduler, tc scheduler) from a saved model."""
    with open(path) as f:
        model = json.load(f)
    out = []
    for key in ("bfs", "tc"):
        s = GpuScheduler(device)
        s.set(params=model[key]["params"], mean=model[key]["mean"], scale=model[key]["scale"])
        out.append(s)
    return tuple(out)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="train the learned schedules on this GPU")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--out", default=MODEL)
    tr.add_argument("--quick", action="store_true", help="a small training set (tests, smoke)")
    tr.add_argument("--steps", type=int, default=3000)
    tr.add_argument("--bfs-targets", choices=["centered", "absolute"], default="centered",
                    help="BFS targets: ln-times less their per-level mean (default), or ln-times")
    args = ap.parse_args(argv)
    device = GpuDevice(0)
    t0 = time.time()
    data = collect(device, training_graphs(args.quick))
    info = {"device": device.name, "collect_seconds": time.time() - t0, "steps": args.steps,
            "training_set": data["names"], "bfs_targets": args.bfs_targets}
    result = {}
    for key in ("bfs", "tc"):
        x, t, mask = data[key]
        # BFS decisions compare predictions only with each other (and a margin in ln-time);
        # triangle-count decisions also weigh a predicted saving in microseconds
        if key == "bfs" and args.bfs_targets == "centered":
            t = centered(t, mask)
        s, before, after = fit(device, x, t, mask, steps=args.steps)
        info[key] = {"samples": int(len(x)), "loss_before": before, "loss_after": after}
        result[key] = s
        _log(f"{key}: {len(x)} samples, loss {before:.4f} -> {after:.4f}")
    save(args.out, result["bfs"], result["tc"], info)
    _log(f"saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
