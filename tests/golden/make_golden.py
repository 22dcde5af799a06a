"""Writes the golden fixtures under tests/golden/ (run once here, results committed).

Inputs are the deterministic generators of the library (graphs are identified by their CSR
SHA-256, which the tests recompute). Expected values come from independent library routines
wherever those finish: networkx (triangles, k-truss, shortest paths, PageRank) and scipy
(PageRank fixed point by a sparse linear solve). For the large R-MAT configurations, where
networkx is too slow, the values come from the C oracle's fast variants, which the CPU tests
check against networkx and against the oracle's literal restatements at smaller sizes; each
fixture records its source.

Usage: python tests/golden/make_golden.py [--large]
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys

import networkx as nx
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for p in ("grafeo-cuda_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(REPO, p))

import oracle  # noqa: E402
from graphs import catalog, csr_sha256, er_graph, rmat_graph  # noqa: E402


def nx_graph(off, tgt):
    n = len(off) - 1
    g = nx.Graph()
    g.add_nodes_from(range(n))
    for u in range(n):
        for v in tgt[off[u]:off[u + 1]]:
            g.add_edge(u, int(v))
    return g


def edges_sha(edges) -> str:
    return hashlib.sha256(np.asarray(edges, dtype="<u4").reshape(-1).tobytes()).hexdigest()


def truss_edges(g: nx.Graph, k: int):
    """networkx k-truss (each edge in >= k-2 triangles) on the loop-free graph, as sorted (u<v)."""
    h = g.copy()
    h.remove_edges_from(list(nx.selfloop_edges(h)))
    t = nx.k_truss(h, k)
    return sorted((min(u, v), max(u, v)) for u, v in t.edges())


def pagerank_fixed_point_scipy(off, tgt, d=0.85):
    """Solves r = (1-d)/n + d (P^T r + (D . r)/n) exactly: (I - d M) r = (1-d)/n 1."""
    n = len(off) - 1
    deg = np.diff(off).astype(np.float64)
    rows = np.repeat(np.arange(n), np.diff(off))
    cols = np.asarray(tgt, dtype=np.int64)
    # pull: r'[v] += d * r[u]/deg[u] for u in N(v)
    vals = 1.0 / deg[cols]
    a = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    dangling = (deg == 0).astype(np.float64)
    m = sp.identity(n) - d * a - d * sp.csr_matrix(np.outer(np.ones(n), dangling) / n)
    r = spla.spsolve(sp.csc_matrix(m), np.full(n, (1 - d) / n))
    return r


def make_catalog():
    out = []
    for name, off, tgt in catalog():
        n = len(off) - 1
        g = nx_graph(off, tgt)
        entry = {
            "name": name,
            "n": n,
            "csr_sha256": csr_sha256(off, tgt),
            "triangles": sum(nx.triangles(nx.Graph(g)).values()) // 3,
            "bfs": [],
            "k_truss": {},
            "pagerank_networkx": None,
            "pagerank_scipy": None,
        }
        h = g.copy()
        h.remove_edges_from(list(nx.selfloop_edges(h)))
        entry["triangles"] = sum(nx.triangles(h).values()) // 3
        for s in range(n):
            lengths = nx.single_source_shortest_path_length(g, s)
            entry["bfs"].append([lengths.get(v, -1) for v in range(n)])
        for k in range(3, 9):
            entry["k_truss"][str(k)] = truss_edges(g, k)
        if n:
            pr = nx.pagerank(g.to_directed(), alpha=0.85, tol=1e-15, max_iter=10000)
            entry["pagerank_networkx"] = [pr[v] for v in range(n)]
            entry["pagerank_scipy"] = [float(x) for x in pagerank_fixed_point_scipy(off, tgt)]
        out.append(entry)
    return out


def make_er():
    off, tgt = er_graph(4096, 16, seed=1)
    g = nx_graph(off, tgt)
    return {
        "spec": "Erdos-Renyi n=4096, 32768 draws, SplitMix64 seed 1, symmetrized, deduplicated, no loops",
        "csr_sha256": csr_sha256(off, tgt),
        "num_edges": int(len(tgt)),
        "triangles": sum(nx.triangles(g).values()) // 3,
        "k_truss": {str(k): {"edges": len(e), "sha256": edges_sha(e)} for k in (3, 4) for e in [truss_edges(g, k)]},
        "source": "networkx",
    }


def make_rmat(large: bool):
    out = {}
    scales = [10, 12, 14] + ([16, 18, 20, 22] if large else [])
    for scale in scales:
        off, tgt = rmat_graph(scale, 16, seed=1)
        key = f"rmat{scale}_ef16_seed1"
        e = {"scale": scale, "edge_factor": 16, "seed": 1, "csr_sha256": csr_sha256(off, tgt),
             "num_edges": int(len(tgt))}
        if scale <= 14:
            g = nx_graph(off, tgt)
            e["triangles"] = sum(nx.triangles(g).values()) // 3
            e["source"] = "networkx"
            if scale <= 12:
                kt = truss_edges(g, 4)
                e["k_truss_4"] = {"edges": len(kt), "sha256": edges_sha(kt), "source": "networkx"}
        else:
            oracle.set_threads(os.cpu_count() or 1)
            e["triangles"] = oracle.triangle_count(off, tgt, fast=True)
            e["source"] = "oracle (fast)"
            if scale == 20:
                kt, it = oracle.k_truss(off, tgt, 4, fast=True)
                e["k_truss_4"] = {"edges": len(kt), "iterations": it, "sha256": edges_sha(kt),
                                  "source": "oracle (fast)"}
                deg = np.diff(off.astype(np.int64))
                rng = np.random.default_rng(1)
                sources = [int(s) for s in rng.choice(np.flatnonzero(deg > 0), size=64, replace=False)]
                bfs = []
                for s in sources:
                    dist, reach, depth = oracle.bfs(off, tgt, s)
                    bfs.append({"source": s, "reachable": reach, "max_depth": depth,
                                "sha256": hashlib.sha256(dist.astype("<i4").tobytes()).hexdigest()})
                e["bfs"] = bfs
        out[key] = e
        print(key, {k: v for k, v in e.items() if k not in ("bfs",)}, flush=True)
    return out


def bench_sources(off, count: int, seed: int) -> list[int]:
    """The BFS sources of bench.py's step: `count` distinct seeded random vertices of nonzero
    degree (numpy PCG64 seeded with `seed`, choice without replacement)."""
    deg = np.diff(off.astype(np.int64))
    rng = np.random.default_rng(seed)
    return [int(v) for v in rng.choice(np.flatnonzero(deg > 0), size=count, replace=False)]


def make_rmat22_bfs():
    """R-MAT 22 (the benchmark's graph): the oracle's distances from the step's 8 sources, by
    SHA-256 of the int32 array, with reachable counts and depths."""
    off, tgt = rmat_graph(22, 16, seed=1)
    oracle.set_threads(os.cpu_count() or 1)
    out = {"graph": "rmat22_ef16_seed1", "csr_sha256": csr_sha256(off, tgt), "source": "oracle_bfs",
           "sources": {}}
    for s in bench_sources(off, 8, 1):
        dist, reach, depth = oracle.bfs(off, tgt, s)
        out["sources"][str(s)] = {"reachable": reach, "max_depth": depth,
                                  "sha256": hashlib.sha256(dist.astype("<i4").tobytes()).hexdigest()}
        print("rmat22 bfs", s, reach, depth, flush=True)
    return out


def make_rmat24():
    """R-MAT 24 (the scheduler's evaluation graph): its triangle count from the oracle's fast
    variant and its CSR SHA-256."""
    off, tgt = rmat_graph(24, 16, seed=1)
    oracle.set_threads(os.cpu_count() or 1)
    e = {"scale": 24, "edge_factor": 16, "seed": 1, "csr_sha256": csr_sha256(off, tgt), "num_edges": int(len(tgt)),
         "triangles": oracle.triangle_count(off, tgt, fast=True), "source": "oracle (fast)"}
    print("rmat24", e, flush=True)
    return {"rmat24_ef16_seed1": e}


def make_batch():
    """The 256-graph uneven batch (grafeo_cuda.workloads.uneven_batch, seed 1): per graph its
    size, CSR SHA-256 and triangle count from the C oracle's fast variant (checked against
    networkx and the literal restatement by the CPU tests)."""
    from grafeo_cuda import workloads

    graphs = []
    for off, tgt in workloads.uneven_batch(256, 1):
        graphs.append({"num_nodes": len(off) - 1, "csr_entries": int(len(tgt)),
                       "csr_sha256": csr_sha256(off, tgt),
                       "triangles": int(oracle.triangle_count(off, tgt, fast=True))})
    return {"uneven_batch_256_seed1": {"source": "oracle_triangle_count_fast", "graphs": graphs,
                                       "total_triangles": sum(g["triangles"] for g in graphs)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--large", action="store_true", help="also R-MAT 16..22 (minutes)")
    ap.add_argument("--only-batch", action="store_true", help="only batch_counts.json")
    ap.add_argument("--only-rmat22-bfs", action="store_true", help="only rmat22_bfs.json")
    ap.add_argument("--only-rmat24", action="store_true", help="only R-MAT 24's entry of rmat_counts.json")
    args = ap.parse_args()
    path = os.path.join(HERE, "rmat_counts.json")
    if args.only_rmat22_bfs:
        with open(os.path.join(HERE, "rmat22_bfs.json"), "w") as f:
            json.dump(make_rmat22_bfs(), f, indent=1)
        return
    if args.only_rmat24:
        existing = json.load(open(path)) if os.path.exists(path) else {}
        existing.update(make_rmat24())
        with open(path, "w") as f:
            json.dump(existing, f, indent=1)
        return
    with open(os.path.join(HERE, "batch_counts.json"), "w") as f:
        json.dump(make_batch(), f, indent=1)
    if args.only_batch:
        return
    with open(os.path.join(HERE, "catalog.json"), "w") as f:
        json.dump(make_catalog(), f, indent=1)
    with open(os.path.join(HERE, "er4096.json"), "w") as f:
        json.dump(make_er(), f, indent=1)
    existing = json.load(open(path)) if os.path.exists(path) else {}
    existing.update(make_rmat(args.large))
    with open(path, "w") as f:
        json.dump(existing, f, indent=1)


if __name__ == "__main__":
    main()
