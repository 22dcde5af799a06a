"""The learned scheduler on the GPU.

  - Its forward pass, loss gradient and Adam training kernels against the FP64 CPU oracle
    (oracle/mlp.py, pinned by torch.autograd and finite differences).
  - Any schedule gives the reference's results: every BFS level method and every triangle-count
    method is forced in turn (output biases that make one choice win), and schedulers with random
    weights pick arbitrary mixtures; results stay bit-exact against the oracle.
  - A network trained on the GPU, on timings the kernels produce on small graphs, beats Beamer's
    rule on a high-diameter grid it was not trained on, with identical distances, and stays
    within 1.5x of the hand rule that chains small frontiers (the default fixed rule).
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import grafeo_cuda
import mlp
import oracle
from grafeo_cuda import _native
from graphs import catalog, grid_graph, random_graph, rmat_graph, undirected

pytestmark = pytest.mark.gpu


def batch(seed, count=300):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 2, (count, 12)).astype(np.float32)
    t = rng.normal(0, 1, (count, 3)).astype(np.float32)
    mask = (rng.random((count, 3)) < 0.8).astype(np.float32)
    mask[0] = 1
    return x, t, mask


def test_forward_and_gradient_against_oracle(device):
    s = grafeo_cuda.GpuScheduler(device, seed=3)
    rng = np.random.default_rng(3)
    mean = rng.normal(0, 1, 12).astype(np.float32)
    scale = rng.uniform(0.5, 2, 12).astype(np.float32)
    s.set(mean=mean, scale=scale)
    p = s.get()["params"].astype(np.float64)
    x, t, mask = batch(4)
    y = s.predict(x)
    y_ref, _ = mlp.forward(p, x, mean, scale)
    assert np.allclose(y, y_ref, rtol=1e-4, atol=1e-5)
    loss, grad = s.loss_grad(x, t, mask)
    loss_ref, grad_ref = mlp.loss_grad(p, x, t, mask, mean, scale)
    assert abs(loss - loss_ref) <= 1e-4 * max(1.0, loss_ref)
    err = np.linalg.norm(grad - grad_ref) / np.linalg.norm(grad_ref)
    assert err < 1e-4, err
    assert np.allclose(grad, grad_ref, rtol=1e-3, atol=1e-5 * np.abs(grad_ref).max())


def test_adam_training_against_oracle(device):
    s = grafeo_cuda.GpuScheduler(device, seed=5)
    p0 = s.get()["params"].astype(np.float64)
    x, t, mask = batch(6, count=200)
    final = s.train(x, t, mask, steps=25, lr=1e-2)
    p_ref, final_ref = mlp.adam(p0, x, t, mask, steps=25, lr=1e-2)
    assert np.allclose(s.get()["params"], p_ref, rtol=1e-3, atol=2e-4)
    assert abs(final - final_ref) <= 1e-3 * max(1.0, final_ref)
    before = mlp.loss(p0, x, t, mask)
    longer = s.train(x, t, mask, steps=300, lr=1e-2)
    assert longer < 0.7 * before


def forced(device, choice):
    """A scheduler whose outputs make `choice` the cheapest wherever it is allowed."""
    s = grafeo_cuda.GpuScheduler(device, seed=1)
    p = np.zeros(1571, np.float32)
    # ln-microsecond outputs 10 apart: far above the triangle count's 20 us switching margin
    b3 = np.full(3, 10.0, np.float32)
    b3[choice] = 0.0
    p[-3:] = b3
    s.set(params=p)
    return s


def check_graph(name, off, tgt, g, sources):
    for source in sources:
        dist, reach, depth = oracle.bfs(off, tgt, source)
        r = g.bfs_numpy(source)
        assert np.array_equal(r["distances"], dist), f"{name}: BFS {source}"
        assert (r["reachable"], r["max_depth"]) == (reach, depth), f"{name}: BFS {source}"
    assert g.triangle_count()["total_triangles"] == oracle.triangle_count(off, tgt), name


def graphs():
    out = [(name, off, tgt) for name, off, tgt in catalog()]
    rng = np.random.default_rng(17)
    for i in range(12):
        n, edges = random_graph(rng, max_nodes=30, max_edges=90)
        out.append((f"random #{i}", *undirected(n, edges)))
    out.append(("RMAT-12", *rmat_graph(12, 16, seed=3)))
    out.append(("grid 40x30", *grid_graph(40, 30)))
    off, tgt = _native.from_edge_pairs([(0, 1), (1, 2), (2, 0), (2, 3), (4, 2)], 6, False)
    out.append(("directed", off, tgt))
    return out


def test_every_forced_choice_is_bit_exact(device):
    for choice in range(3):
        s = forced(device, choice)
        for name, off, tgt in graphs():
            g = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
            g.set_scheduler(s, s)
            n = len(off) - 1
            check_graph(f"{name} (forced {choice})", off, tgt, g, sorted({0, n // 2, n - 1}) if n else [])
            if n and len(tgt):
                g.bfs_numpy(0)
                sched = g.last_schedule()
                assert sched and all(c in (0, 1, 2) for c in sched)


def test_random_networks_are_bit_exact(device):
    for seed in range(4):
        s = grafeo_cuda.GpuScheduler(device, seed=100 + seed)
        rng = np.random.default_rng(seed)
        s.set(mean=rng.normal(0, 3, 12), scale=rng.uniform(0.2, 3, 12))
        for name, off, tgt in graphs()[-3:]:
            g = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
            g.set_scheduler(s, s)
            n = len(off) - 1
            check_graph(f"{name} (random net {seed})", off, tgt, g, [0, n - 1])


def test_detaching_restores_the_fixed_rule(device, monkeypatch):
    off, tgt = rmat_graph(12, 16, seed=3)
    g = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
    fresh = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
    fresh.bfs_numpy(0)
    g.set_scheduler(bfs=forced(device, 2))
    g.bfs_numpy(0)
    assert 2 in g.last_schedule()
    g.set_scheduler()
    g.bfs_numpy(0)
    assert g.last_schedule() == fresh.last_schedule()
    # Beamer's rule alone never chains
    monkeypatch.setenv("GC_BFS_RULE", "beamer")
    g.bfs_numpy(0)
    assert 2 not in g.last_schedule()
    monkeypatch.delenv("GC_BFS_RULE")


def collect(device, graphs_and_sources):
    xs, ts, ms = [], [], []
    for off, tgt, sources in graphs_and_sources:
        g = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
        for s in sources:
            r = g.bfs_level_timings(s)
            t = r["times_us"].astype(np.float64)
            xs.append(r["features"])
            ts.append(np.log(np.maximum(t, 1.0)))
            ms.append((t > 0).astype(np.float32))
    return np.concatenate(xs), np.concatenate(ts).astype(np.float32), np.concatenate(ms)


def test_trained_network_beats_the_fixed_rule_on_a_grid(device):
    rng = np.random.default_rng(0)
    train = []
    for rows in (64, 128, 256, 512):
        off, tgt = grid_graph(rows, rows)
        n = len(off) - 1
        train.append((off, tgt, [int(v) for v in rng.integers(0, n, 2)]))
    for scale in (10, 12, 14):
        off, tgt = rmat_graph(scale, 16, seed=scale)
        deg = np.diff(off.astype(np.int64))
        train.append((off, tgt, [int(v) for v in rng.choice(np.flatnonzero(deg > 0), 2)]))
    x, t, mask = collect(device, train)
    assert len(x) > 100
    s = grafeo_cuda.GpuScheduler(device, seed=7)
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    s.set(mean=mean, scale=1.0 / np.where(std > 1e-6, std, 1.0))
    before = s.loss_grad(x, t, mask)[0]
    after = s.train(x, t, mask, steps=1500, lr=3e-3)
    assert after < 0.5 * before, (before, after)

    off, tgt = grid_graph(1024, 1024)
    g = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
    source = 1024 * 512 + 512

    def timed():
        first = g.bfs_numpy(source)
        return first, min(g.bfs_numpy(source)["time_seconds"] for _ in range(3))

    os.environ["GC_BFS_RULE"] = "beamer"
    try:
        beamer, beamer_t = timed()
    finally:
        del os.environ["GC_BFS_RULE"]
    chain, chain_t = timed()  # the default fixed rule: Beamer with chained top-down levels
    g.set_scheduler(bfs=s)
    learned, learned_t = timed()
    assert np.array_equal(learned["distances"], beamer["distances"])
    assert np.array_equal(chain["distances"], beamer["distances"])
    r, c = np.divmod(np.arange(1024 * 1024), 1024)
    assert np.array_equal(beamer["distances"], np.abs(r - 512) + np.abs(c - 512))
    print(f"grid 1024x1024 BFS: Beamer {beamer_t * 1e3:.1f} ms, Beamer + chain {chain_t * 1e3:.1f} ms, "
          f"learned {learned_t * 1e3:.1f} ms")
    # Trained on timings alone, the network finds the chained levels on a graph it never saw
    # (Beamer's rule has no such choice); the hand rule that chains small frontiers is the bar
    # it is compared with in the benchmark, where it does not beat it.
    assert learned_t < 0.8 * beamer_t, (beamer_t, learned_t)
    assert learned_t < 1.5 * chain_t, (chain_t, learned_t)
