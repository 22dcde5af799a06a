"""Summaries of rocprofv3 CSV output for the library's kernels.

    python3 -m grafeo_cuda.profiling stats DIR
        Markdown table of DIR's kernel statistics (`rocprofv3 --kernel-trace --stats
        --output-format csv -d DIR`): calls, total, average, min and max duration per kernel.

    python3 -m grafeo_cuda.profiling traffic --workload KEY --fetch DIR --write DIR \\
            --copy-fetch DIR --copy-write DIR [--copy-bytes N] [--out FILE]
        HBM bytes per launch of every kernel from two counter runs (`--pmc FETCH_SIZE`, then
        `--pmc WRITE_SIZE`, KiB per dispatch), each scaled by what the same counter reads for a
        plain copy of known size (N bytes read and N written per launch). Writes JSON.

Only the standard library is used; nothing here touches the GPU.
"""

from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
from collections import defaultdict


def _rows(directory: str, suffix: str) -> list[dict]:
    """Rows of every `*<suffix>` CSV under `directory`, with lower-cased column names."""
    rows = []
    for path in sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True)):
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                rows.append({k.strip().lower(): v for k, v in row.items() if k is not None})
    if not rows:
        raise SystemExit(f"no *{suffix} under {directory}")
    return rows


def kernel_name(raw: str) -> str:
    name = raw.strip()
    return name[:-3] if name.endswith(".kd") else name


def stats_table(directory: str, top: int = 25) -> str:
    rows = _rows(directory, "kernel_stats.csv")
    rows.sort(key=lambda r: -float(r.get("totaldurationns", 0)))
    out = ["| kernel | calls | total ms | avg us | min us | max us | % |",
           "|---|---:|---:|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append("| {} | {} | {:.3f} | {:.1f} | {:.1f} | {:.1f} | {:.1f} |".format(
            kernel_name(r["name"]), r["calls"], float(r["totaldurationns"]) / 1e6,
            float(r["averagens"]) / 1e3, float(r["minns"]) / 1e3, float(r["maxns"]) / 1e3,
            float(r.get("percentage", 0.0))))
    return "\n".join(out)


def counter_means(directory: str, counter: str) -> dict[str, tuple[float, int]]:
    """Mean counter value per dispatch and dispatch count, per kernel."""
    per_dispatch: dict[tuple[str, str], float] = defaultdict(float)
    for r in _rows(directory, "counter_collection.csv"):
        if r.get("counter_name", "").strip() != counter:
            continue
        key = (kernel_name(r["kernel_name"]), r.get("dispatch_id", r.get("correlation_id", "")))
        per_dispatch[key] += float(r["counter_value"])
    sums: dict[str, list[float]] = defaultdict(list)
    for (name, _), v in per_dispatch.items():
        sums[name].append(v)
    return {name: (sum(v) / len(v), len(v)) for name, v in sums.items()}


def traffic(workload: str, fetch_dir: str, write_dir: str, copy_fetch_dir: str,
            copy_write_dir: str, copy_bytes: int) -> dict:
    fetch = counter_means(fetch_dir, "FETCH_SIZE")
    write = counter_means(write_dir, "WRITE_SIZE")
    copy_fetch = counter_means(copy_fetch_dir, "FETCH_SIZE")
    copy_write = counter_means(copy_write_dir, "WRITE_SIZE")

    def largest(d):
        return max(d.items(), key=lambda kv: kv[1][0])

    copy_name, (cf_kib, _) = largest(copy_fetch)
    _, (cw_kib, _) = largest(copy_write)
    fetch_factor = copy_bytes / 1024.0 / cf_kib
    write_factor = copy_bytes / 1024.0 / cw_kib
    kernels = {}
    for name in sorted(set(fetch) & set(write)):
        f_kib, calls = fetch[name]
        w_kib, _ = write[name]
        kernels[name] = {
            "dispatches": calls,
            "fetch_kib_per_dispatch": f_kib,
            "write_kib_per_dispatch": w_kib,
            "hbm_bytes_per_launch": 1024.0 * (f_kib * fetch_factor + w_kib * write_factor),
        }
    return {
        "workload": workload,
        "copy": {"kernel": copy_name, "bytes_read": copy_bytes, "bytes_written": copy_bytes,
                 "fetch_kib": cf_kib, "write_kib": cw_kib},
        "fetch_factor": fetch_factor,
        "write_factor": write_factor,
        "kernels": kernels,
    }


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=25)
    t = sub.add_parser("traffic")
    t.add_argument("--workload", required=True)
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-bytes", type=int, default=1 << 30)
    t.add_argument("--out")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(stats_table(args.dir, args.top))
        return 0
    result = traffic(args.workload, args.fetch, args.write, args.copy_fetch, args.copy_write,
                     args.copy_bytes)
    text = json.dumps(result, indent=1, sort_keys=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    else:
        print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
