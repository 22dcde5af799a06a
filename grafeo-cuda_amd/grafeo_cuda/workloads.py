"""Seeded synthetic workloads built from the library's generators (gc_gen_*).

uneven_batch(count, seed): the batched triangle-count workload. Graph i (0-based):
  - n_i = round(exp(U_i)), U_i uniform on [ln 1024, ln 65536] (log-uniform vertex count);
  - even i: Erdos-Renyi draws, n_i * d_i / 2 pairs, average degree d_i uniform on [8, 32];
  - odd i: R-MAT, scale s_i = ceil(log2 n_i), edge factor 8, (a, b, c) = (0.57, 0.19, 0.19),
    vertex ids permuted; draws with an endpoint >= n_i are dropped (ids are kept, not folded);
  - each graph is symmetrized, deduplicated and stripped of self-loops into a u32 CSR.
The random draws (n_i, d_i) come from numpy's PCG64 generator seeded with `seed`; graph i's
generator seed is seed * 1_000_003 + i.
"""

from __future__ import annotations

import math

import numpy as np

from . import _native


def uneven_batch(count: int = 256, seed: int = 1, min_nodes: int = 1024, max_nodes: int = 65536):
    """List of (offsets, targets) u32 CSRs, see the module docstring."""
    rng = np.random.default_rng(seed)
    graphs = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(min_nodes), math.log(max_nodes)))))
        n = max(min_nodes, min(max_nodes, n))
        gseed = seed * 1_000_003 + i
        if i % 2 == 0:
            degree = rng.uniform(8.0, 32.0)
            pairs = _native.gen_erdos_renyi(n, int(n * degree / 2), gseed)
        else:
            scale = max(1, math.ceil(math.log2(n)))
            pairs = _native.gen_rmat(scale, 8, gseed)
            pairs = pairs[(pairs[:, 0] < n) & (pairs[:, 1] < n)]
        graphs.append(_native.from_edge_pairs(pairs, n, True))
    return graphs
