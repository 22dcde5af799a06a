"""Named small graphs and generated inputs shared by the tests.

The catalog mirrors the reference's test graphs (crates/grafeo-cuda-algos/tests/correctness.rs:93-210
and tests/roundtrip.rs:9-43), built as symmetric CSR the way its `undirected()` helper plus
`export_to_csr(store, true)` does: each edge in both directions, a self-loop once, parallel
edges collapsed.
"""

from __future__ import annotations

import hashlib

import numpy as np

from grafeo_cuda import _native


def complete_edges(n: int):
    return [(i, j) for i in range(n) for j in range(i + 1, n)]


def path_edges(n: int):
    return [(i - 1, i) for i in range(1, n)]


PETERSEN = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 5), (1, 6), (2, 7), (3, 8), (4, 9),
            (5, 7), (7, 9), (9, 6), (6, 8), (8, 5)]


def _catalog_specs():
    star = [(0, leaf) for leaf in range(1, 9)]
    wheel = [e for i in range(1, 7) for e in ((0, i), (i, i % 6 + 1))]
    two_k4 = complete_edges(4) + [(a + 4, b + 4) for a, b in complete_edges(4)] + [(3, 4)]
    k4_tail = complete_edges(4) + [(3, 4), (4, 5)]
    # "messy" (correctness.rs:177-210): node 2 of 7 deleted; live nodes 0,1,3,4,5,6 become
    # 0..5; edges touching node 2 disappear; self-loops at 0 and 4; parallel edges collapse.
    live = {0: 0, 1: 1, 3: 2, 4: 3, 5: 4, 6: 5}
    raw = [(0, 1), (1, 3), (0, 3), (3, 4), (4, 5), (5, 3), (1, 2), (2, 6), (6, 0), (5, 6), (4, 6),
           (0, 0), (4, 4), (0, 1), (4, 3), (5, 6), (6, 5)]
    messy = [(live[a], live[b]) for a, b in raw if a in live and b in live]
    return [
        ("single node", 1, []),
        ("edgeless", 4, []),
        ("K4", 4, complete_edges(4)),
        ("K5", 5, complete_edges(5)),
        ("K7", 7, complete_edges(7)),
        ("path", 10, path_edges(10)),
        ("petersen", 10, PETERSEN),
        ("star", 9, star),
        ("wheel", 7, wheel),
        ("diamond", 4, [(0, 1), (0, 2), (1, 2), (1, 3), (2, 3)]),
        ("bowtie", 5, [(0, 1), (0, 2), (1, 2), (0, 3), (0, 4), (3, 4)]),
        ("triangle, edge and isolated nodes", 7, [(0, 1), (1, 2), (0, 2), (3, 4)]),
        ("two K4 joined by a bridge", 8, two_k4),
        ("K4 with a tail", 6, k4_tail),
        ("messy", 6, messy),
    ]


def catalog():
    """[(name, offsets, targets)] as symmetric CSR (numpy uint32)."""
    out = []
    for name, n, edges in _catalog_specs():
        off, tgt = _native.from_edge_pairs(edges, n, True)
        out.append((name, off, tgt))
    return out


def undirected(n: int, edges):
    return _native.from_edge_pairs(edges, n, True)


def random_graph(rng: np.random.Generator, max_nodes=16, max_edges=48, loops=True):
    """The reference's proptest shape (correctness.rs:483-513): n in [1, 16), up to 48 pairs."""
    n = int(rng.integers(1, max_nodes))
    e = int(rng.integers(0, max_edges + 1))
    pairs = rng.integers(0, n, size=(e, 2))
    if not loops:
        pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    return n, [tuple(map(int, p)) for p in pairs]


def csr_sha256(offsets: np.ndarray, targets: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(offsets, dtype="<u4").tobytes())
    h.update(np.ascontiguousarray(targets, dtype="<u4").tobytes())
    return h.hexdigest()


def array_sha256(a: np.ndarray, dtype: str) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def er_graph(n: int, avg_degree: int, seed: int):
    """Erdos-Renyi input: n*avg_degree/2 uniform draws, symmetrized, deduplicated, no loops."""
    pairs = _native.gen_erdos_renyi(n, n * avg_degree // 2, seed)
    return _native.from_edge_pairs(pairs, n, True)


def rmat_graph(scale: int, edge_factor: int, seed: int):
    """R-MAT (0.57, 0.19, 0.19, 0.05) with permuted ids, symmetrized, deduplicated, no loops."""
    pairs = _native.gen_rmat(scale, edge_factor, seed)
    return _native.from_edge_pairs(pairs, 1 << scale, True)


def grid_graph(rows: int, cols: int):
    return _native.from_edge_pairs(_native.gen_grid(rows, cols), rows * cols, True)
