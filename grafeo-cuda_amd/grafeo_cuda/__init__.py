"""grafeo_cuda - the Python surface of grafeo-cuda, on MI355X (gfx950).

Drop-in for the reference's PyO3 module `grafeo_cuda` (crates/bindings/python/src/lib.rs:12-17,
crates/bindings/python/src/gpu_graph.rs:11-155): the classes `GpuDevice` and `GpuGraph`, with
the same constructors, properties, method names, keyword defaults, result dict keys, and
`RuntimeError(str(CudaError))` on failure. Everything runs through the C ABI of
libgrafeo_hip.so (include/grafeo_hip.h); there is no CPU fallback.

Supersets of the reference surface: `from_csr` also accepts numpy arrays (no per-element
conversion), `triangle_count_oriented`, `prepare`, `is_symmetric`, `msbfs` (many BFS sources in
one traversal), `truss_numbers`, `GpuBatch` (many graphs triangle-counted in one launch), and
array-returning variants (`*_numpy`) for large graphs. The GIL is released during device work.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_float, c_int32, c_size_t, c_uint32
from ctypes import c_uint64, c_void_p

import numpy as np

from . import _native
from ._native import GcError, check, lib

__all__ = ["GpuDevice", "GpuGraph", "GpuBatch", "GpuScheduler", "GcError"]


def _runtime_error(e: GcError) -> RuntimeError:
    err = RuntimeError(str(e))
    err.status = e.status
    err.kind = e.kind
    return err


def _u32_arg(value, what: str) -> int:
    if isinstance(value, bool) or not isinstance(value, (int, np.integer)):
        raise TypeError(f"{what} must be an int")
    if value < 0 or value > 0xFFFFFFFF:
        raise OverflowError(f"{what} out of range for u32")
    return int(value)


class GpuDevice:
    """A GPU device (GpuDevice, crates/bindings/python/src/gpu_graph.rs:11-37)."""

    def __init__(self, ordinal: int = 0):
        ordinal = _u32_arg(ordinal, "ordinal")
        handle = c_void_p()
        try:
            check(lib.gc_device_new(ordinal, byref(handle)))
        except GcError as e:
            raise _runtime_error(e) from None
        self._handle = handle
        name = ctypes.create_string_buffer(256)
        total, ordn = c_uint64(), c_uint32()
        check(lib.gc_device_properties(handle, name, 256, byref(total), byref(ordn)))
        self._name = name.value.decode(errors="replace")
        self._total_memory = total.value
        self._ordinal = ordn.value

    @property
    def name(self) -> str:
        """Device name."""
        return self._name

    @property
    def total_memory(self) -> int:
        """Total GPU memory in bytes."""
        return self._total_memory

    @property
    def ordinal(self) -> int:
        return self._ordinal

    def profile(self, on: bool = True, only=None) -> None:
        """Starts (or stops) per-kernel event timing and clears accumulated figures; `only`
        (kernel names) limits the timing to those kernels, None times every kernel."""
        names = ",".join(only) if only else ""
        check(lib.gc_profile_only(self._handle, names.encode()))
        check(lib.gc_profile_enable(self._handle, 1 if on else 0))

    def kernel_time(self, kernel: str) -> tuple[float, int]:
        """(total milliseconds, launches) recorded for `kernel` since `profile()`."""
        ms, count = c_double(), c_uint64()
        check(lib.gc_profile_query(self._handle, kernel.encode(), byref(ms), byref(count)))
        return ms.value, count.value

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            lib.gc_device_release(h)
            self._handle = None

    def __repr__(self) -> str:
        return f"GpuDevice(ordinal={self._ordinal}, name={self._name!r})"


class GpuGraph:
    """A graph in GPU memory (GpuGraph, crates/bindings/python/src/gpu_graph.rs:39-155)."""

    def __init__(self, handle: c_void_p, device: GpuDevice):
        self._handle = handle
        self._device = device  # keeps the Python device object alive too

    @staticmethod
    def from_csr(offsets, targets, device: GpuDevice) -> "GpuGraph":
        """Create a GpuGraph from CSR arrays (lists of ints or numpy arrays)."""
        if not isinstance(device, GpuDevice):
            raise TypeError("device must be a GpuDevice")
        off = _native.as_u32(offsets, "offsets")
        tgt = _native.as_u32(targets, "targets")
        handle = c_void_p()
        try:
            check(
                lib.gc_graph_from_csr(
                    device._handle, _native.u32_ptr(off), off.size, _native.u32_ptr(tgt), tgt.size,
                    byref(handle),
                )
            )
        except GcError as e:
            raise _runtime_error(e) from None
        return GpuGraph(handle, device)

    @staticmethod
    def from_edge_pairs(edges, num_nodes: int, device: GpuDevice, symmetric: bool = True) -> "GpuGraph":
        """A GpuGraph whose CSR is built on the GPU from (source, target) pairs (list or (E, 2)
        array): from_edge_pairs' arrays without a host CSR."""
        if not isinstance(device, GpuDevice):
            raise TypeError("device must be a GpuDevice")
        pairs = _native.pairs_array(edges)
        handle = c_void_p()
        try:
            check(lib.gc_graph_from_edge_pairs(device._handle, _native.u32_ptr(pairs), pairs.shape[0],
                                               _u32_arg(num_nodes, "num_nodes"), int(bool(symmetric)),
                                               byref(handle)))
        except GcError as e:
            raise _runtime_error(e) from None
        return GpuGraph(handle, device)

    @staticmethod
    def from_tsv(path: str, device: GpuDevice, symmetric: bool = True) -> "GpuGraph":
        """A GpuGraph from a TSV edge list parsed and built on the GPU: parse_tsv_to_edges' pairs
        (ids in order of first appearance), then from_edge_pairs' CSR, with no host CSR. The number
        of edge lines parsed is in `.parsed_pairs`."""
        if not isinstance(device, GpuDevice):
            raise TypeError("device must be a GpuDevice")
        handle, count = c_void_p(), c_size_t()
        try:
            check(lib.gc_graph_from_tsv(device._handle, os.fsencode(path), int(bool(symmetric)), byref(handle),
                                        byref(count)))
        except GcError as e:
            raise _runtime_error(e) from None
        g = GpuGraph(handle, device)
        g.parsed_pairs = count.value
        return g

    @property
    def num_nodes(self) -> int:
        """Number of nodes."""
        return lib.gc_graph_num_nodes(self._handle)

    @property
    def num_edges(self) -> int:
        """Number of edge entries in CSR."""
        return lib.gc_graph_num_edges(self._handle)

    def _call(self, fn, *args) -> None:
        try:
            check(fn(self._handle, *args))
        except GcError as e:
            raise _runtime_error(e) from None

    def prepare(self) -> float:
        """Builds the derived device layouts now; returns the seconds it took."""
        secs = c_double()
        self._call(lib.gc_graph_prepare, byref(secs))
        return secs.value

    def relabelled_numpy(self) -> dict:
        """The degree-ordered layout the kernels run on, copied to the host: {'offsets',
        'targets', 'perm' (new id -> id), 'fwd' (first entry above each row)}."""
        ptrs = [c_void_p() for _ in range(4)]
        self._call(lib.gc_graph_relabelled, *[byref(p) for p in ptrs])
        n, m = self.num_nodes, self.num_edges
        out = {}
        for name, p, size in zip(("offsets", "targets", "perm", "fwd"), ptrs, (n + 1, m, n, n)):
            arr = np.zeros(max(size, 1), dtype=np.uint32)
            if size:
                check(lib.gc_device_copy_to_host(self._device._handle, arr.ctypes.data_as(c_void_p), p, 4 * size))
            out[name] = arr[:size]
        return out

    def is_symmetric(self) -> bool:
        sym = c_int32()
        self._call(lib.gc_graph_is_symmetric, byref(sym))
        return bool(sym.value)

    def triangle_count(self) -> dict:
        """Count all triangles in the graph on GPU: {'total_triangles', 'time_seconds'}."""
        total, secs = c_uint64(), c_double()
        self._call(lib.gc_triangle_count, byref(total), byref(secs))
        return {"total_triangles": total.value, "time_seconds": secs.value}

    def triangle_count_oriented(self) -> dict:
        """triangle_counting::triangle_count_oriented for a degree-oriented CSR."""
        total, secs = c_uint64(), c_double()
        self._call(lib.gc_triangle_count_oriented, byref(total), byref(secs))
        return {"total_triangles": total.value, "time_seconds": secs.value}

    def pagerank_numpy(self, damping: float = 0.85, max_iterations: int = 100, tolerance: float = 1e-6) -> dict:
        """As pagerank, with 'ranks' as a float32 numpy array."""
        max_iterations = _u32_arg(max_iterations, "max_iterations")
        ranks = np.zeros(self.num_nodes, dtype=np.float32)
        it, conv, secs = c_uint32(), c_int32(), c_double()
        self._call(
            lib.gc_pagerank, c_float(damping), max_iterations, c_float(tolerance),
            ranks.ctypes.data_as(POINTER(c_float)), byref(it), byref(conv), byref(secs),
        )
        return {"ranks": ranks, "iterations": it.value, "converged": bool(conv.value), "time_seconds": secs.value}

    def pagerank(self, damping: float = 0.85, max_iterations: int = 100, tolerance: float = 1e-6) -> dict:
        """Run PageRank on GPU: {'ranks', 'iterations', 'converged', 'time_seconds'}."""
        r = self.pagerank_numpy(damping, max_iterations, tolerance)
        r["ranks"] = [float(x) for x in r["ranks"]]
        return r

    def bfs_numpy(self, source: int) -> dict:
        """As bfs, with 'distances' as an int32 numpy array."""
        source = _u32_arg(source, "source")
        dist = np.zeros(self.num_nodes, dtype=np.int32)
        reach, depth, secs = c_uint32(), c_int32(), c_double()
        self._call(
            lib.gc_bfs, source, dist.ctypes.data_as(POINTER(c_int32)), byref(reach), byref(depth), byref(secs)
        )
        return {"distances": dist, "reachable": reach.value, "max_depth": depth.value, "time_seconds": secs.value}

    def bfs(self, source: int) -> dict:
        """Run BFS from a source node: {'distances', 'reachable', 'max_depth', 'time_seconds'}."""
        r = self.bfs_numpy(source)
        r["distances"] = [int(x) for x in r["distances"]]
        return r

    def msbfs_numpy(self, sources) -> dict:
        """BFS from every source in one call (128 sources per traversal): {'distances':
        int32 array (len(sources), num_nodes), 'reachable': uint32 array, 'max_depth': int32
        array, 'time_seconds'}; row s is what bfs(sources[s]) returns."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int64))
        if src.ndim != 1:
            raise TypeError("sources must be a sequence of node ids")
        if src.size and (src.min() < 0 or src.max() > 0xFFFFFFFF):
            raise OverflowError("source ids must fit in u32")
        src = src.astype(np.uint32)
        count = int(src.size)
        dist = np.zeros((count, self.num_nodes), dtype=np.int32)
        reach = np.zeros(count, dtype=np.uint32)
        depth = np.zeros(count, dtype=np.int32)
        secs = c_double()
        self._call(
            lib.gc_msbfs, src.ctypes.data_as(POINTER(c_uint32)), count,
            dist.ctypes.data_as(POINTER(c_int32)), reach.ctypes.data_as(POINTER(c_uint32)),
            depth.ctypes.data_as(POINTER(c_int32)), byref(secs),
        )
        return {"distances": dist, "reachable": reach, "max_depth": depth, "time_seconds": secs.value}

    def msbfs(self, sources) -> dict:
        """As msbfs_numpy with lists: {'distances': list of lists, 'reachable', 'max_depth',
        'time_seconds'}."""
        r = self.msbfs_numpy(sources)
        r["distances"] = [[int(x) for x in row] for row in r["distances"]]
        r["reachable"] = [int(x) for x in r["reachable"]]
        r["max_depth"] = [int(x) for x in r["max_depth"]]
        return r

    def k_truss_numpy(self, k: int) -> dict:
        """As k_truss, with 'edges' as an (E, 2) uint32 numpy array."""
        k = _u32_arg(k, "k")
        ptr, count, it, secs = POINTER(c_uint32)(), c_size_t(), c_uint32(), c_double()
        self._call(lib.gc_k_truss, k, byref(ptr), byref(count), byref(it), byref(secs))
        edges = _native.take_u32(ptr, 2 * count.value).reshape(-1, 2)
        return {"edges": edges, "k": k, "iterations": it.value, "time_seconds": secs.value}

    def k_truss(self, k: int) -> dict:
        """Extract the k-truss: {'edges': list[tuple], 'k', 'iterations', 'time_seconds'}."""
        r = self.k_truss_numpy(k)
        r["edges"] = [(int(u), int(v)) for u, v in r["edges"]]
        return r

    def set_scheduler(self, bfs=None, tc=None) -> None:
        """Attaches GpuSchedulers: `bfs` picks each BFS level's method, `tc` each degree bucket's
        triangle-count method; None keeps the fixed rule. Results do not change."""
        self._call(lib.gc_graph_set_scheduler, bfs._handle if bfs is not None else None,
                   tc._handle if tc is not None else None)
        self._schedulers = (bfs, tc)

    def last_schedule(self) -> list[int]:
        """Choices of the last BFS (per level run: 0 top-down, 1 bottom-up, 2 device-chained levels)
        or symmetric triangle count (per degree bucket: 0 bitmap, 1 hash, 2 binary search)."""
        count = c_uint32()
        self._call(lib.gc_graph_last_schedule, None, 0, byref(count))
        out = np.zeros(max(count.value, 1), dtype=np.uint32)
        self._call(lib.gc_graph_last_schedule, out.ctypes.data_as(POINTER(c_uint32)), count.value, byref(count))
        return [int(x) for x in out[: count.value]]

    def bfs_level_timings(self, source: int, max_levels: int = 100000) -> dict:
        """Training data for the BFS decision: {'features': (L, 12), 'times_us': (L, 3), -1 where
        a choice is not allowed}; each level's choices are all timed from the same state."""
        source = _u32_arg(source, "source")
        cap = max(1, min(int(max_levels), self.num_nodes + 1))
        feats = np.zeros((cap, 12), dtype=np.float32)
        times = np.zeros((cap, 3), dtype=np.float32)
        levels = c_uint32()
        self._call(lib.gc_bfs_level_timings, source, cap, feats.ctypes.data_as(POINTER(c_float)),
                   times.ctypes.data_as(POINTER(c_float)), byref(levels))
        return {"features": feats[: levels.value], "times_us": times[: levels.value]}

    def tc_bucket_timings(self, max_buckets: int = 64) -> dict:
        """Training data for the triangle-count decision: per degree bucket {'features': (B, 12),
        'times_us': (B, 3)} (bitmap, hash, binary search, each timed alone on the bucket)."""
        feats = np.zeros((max_buckets, 12), dtype=np.float32)
        times = np.zeros((max_buckets, 3), dtype=np.float32)
        count = c_uint32()
        self._call(lib.gc_tc_bucket_timings, max_buckets, feats.ctypes.data_as(POINTER(c_float)),
                   times.ctypes.data_as(POINTER(c_float)), byref(count))
        return {"features": feats[: count.value], "times_us": times[: count.value]}

    def truss_numbers_numpy(self) -> dict:
        """As truss_numbers, with 'edges' as an (E, 2) uint32 array and 'truss' as a uint32 array."""
        ep, tp = POINTER(c_uint32)(), POINTER(c_uint32)()
        count, top, secs = c_size_t(), c_uint32(), c_double()
        self._call(lib.gc_truss_numbers, byref(ep), byref(tp), byref(count), byref(top), byref(secs))
        edges = _native.take_u32(ep, 2 * count.value).reshape(-1, 2)
        truss = _native.take_u32(tp, count.value)
        return {"edges": edges, "truss": truss, "max_truss": top.value, "time_seconds": secs.value}

    def truss_numbers(self) -> dict:
        """Trussness of every edge: {'edges': list[tuple], 'truss': list[int], 'max_truss',
        'time_seconds'}; k_truss(k)['edges'] lists exactly the edges with truss >= k."""
        r = self.truss_numbers_numpy()
        r["edges"] = [(int(u), int(v)) for u, v in r["edges"]]
        r["truss"] = [int(t) for t in r["truss"]]
        return r

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            lib.gc_graph_free(h)
            self._handle = None

    def __repr__(self) -> str:
        return f"GpuGraph(num_nodes={self.num_nodes}, num_edges={self.num_edges})"


class GpuBatch:
    """Many symmetric graphs triangle-counted together in one launch (gc_batch_*; an addition:
    counts[i] equals GpuGraph.from_csr(*graphs[i]).triangle_count()['total_triangles'])."""

    def __init__(self, handle: c_void_p, device: GpuDevice, keep):
        self._handle = handle
        self._device = device
        self._keep = keep

    @staticmethod
    def from_csrs(graphs, device: GpuDevice) -> "GpuBatch":
        """`graphs`: a sequence of (offsets, targets) pairs, each a symmetric CSR."""
        if not isinstance(device, GpuDevice):
            raise TypeError("device must be a GpuDevice")
        arrays = [(_native.as_u32(o, "offsets"), _native.as_u32(t, "targets")) for o, t in graphs]
        count = len(arrays)
        for off, _ in arrays:
            if off.size == 0:
                raise RuntimeError("invalid CSR: offsets must have at least one element")
        off_ptrs = (POINTER(c_uint32) * max(count, 1))(*[_native.u32_ptr(o) for o, _ in arrays])
        tgt_ptrs = (POINTER(c_uint32) * max(count, 1))(*[_native.u32_ptr(t) for _, t in arrays])
        nodes = np.asarray([o.size - 1 for o, _ in arrays] or [0], dtype=np.uint32)
        sizes = np.asarray([t.size for _, t in arrays] or [0], dtype=np.uint64)
        handle = c_void_p()
        try:
            check(
                lib.gc_batch_from_csrs(
                    device._handle, count, off_ptrs, nodes.ctypes.data_as(POINTER(c_uint32)), tgt_ptrs,
                    sizes.ctypes.data_as(POINTER(c_uint64)), byref(handle),
                )
            )
        except GcError as e:
            raise _runtime_error(e) from None
        return GpuBatch(handle, device, None)

    @property
    def num_graphs(self) -> int:
        return lib.gc_batch_num_graphs(self._handle)

    def _call(self, fn, *args) -> None:
        try:
            check(fn(self._handle, *args))
        except GcError as e:
            raise _runtime_error(e) from None

    def triangle_count(self) -> dict:
        """{'counts': uint64 array (one per graph), 'time_seconds'}."""
        counts = np.zeros(max(self.num_graphs, 1), dtype=np.uint64)
        secs = c_double()
        self._call(lib.gc_batch_triangle_count, counts.ctypes.data_as(POINTER(c_uint64)), byref(secs))
        return {"counts": counts[: self.num_graphs], "time_seconds": secs.value}

    @property
    def config(self) -> tuple:
        """(bitmap out-degree threshold, DAG edges per merge task, rows per bitmap task,
        workgroups per CU)."""
        c = (c_uint32 * 4)()
        self._call(lib.gc_batch_get_config, c)
        return tuple(int(x) for x in c)

    def set_config(self, config) -> None:
        c = (c_uint32 * 4)(*[int(x) for x in config])
        self._call(lib.gc_batch_set_config, c)

    def tune(self) -> dict:
        """Times candidate launch shapes on this batch and keeps the fastest:
        {'config', 'best_ms'}."""
        c = (c_uint32 * 4)()
        ms = c_double()
        self._call(lib.gc_batch_tune, c, byref(ms))
        return {"config": tuple(int(x) for x in c), "best_ms": ms.value}

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            lib.gc_batch_free(h)
            self._handle = None

    def __repr__(self) -> str:
        return f"GpuBatch(num_graphs={self.num_graphs})"


SCHED_PARAMS = 1571


def _f32(a, shape_tail, what):
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if arr.ndim != len(shape_tail) + 1 or tuple(arr.shape[1:]) != tuple(shape_tail):
        raise ValueError(f"{what} must have shape (count, {', '.join(map(str, shape_tail))})")
    return arr


class GpuScheduler:
    """The learned scheduler (gc_sched_*; an addition): a 12 -> 32 -> 32 -> 3 FP32 MLP on the GPU
    predicting the log-time of three choices, with GPU inference, loss gradient and training."""

    def __init__(self, device: GpuDevice, seed: int = 1):
        if not isinstance(device, GpuDevice):
            raise TypeError("device must be a GpuDevice")
        handle = c_void_p()
        try:
            check(lib.gc_sched_new(device._handle, int(seed), byref(handle)))
        except GcError as e:
            raise _runtime_error(e) from None
        self._handle = handle
        self._device = device

    def _call(self, fn, *args) -> None:
        try:
            check(fn(self._handle, *args))
        except GcError as e:
            raise _runtime_error(e) from None

    def get(self) -> dict:
        """{'params': float32[1571], 'mean': float32[12], 'scale': float32[12]}."""
        p = np.zeros(SCHED_PARAMS, np.float32)
        mean = np.zeros(12, np.float32)
        scale = np.zeros(12, np.float32)
        fp = POINTER(c_float)
        self._call(lib.gc_sched_get, p.ctypes.data_as(fp), mean.ctypes.data_as(fp), scale.ctypes.data_as(fp))
        return {"params": p, "mean": mean, "scale": scale}

    def set(self, params=None, mean=None, scale=None) -> None:
        fp = POINTER(c_float)
        arrays = []
        for a, size in ((params, SCHED_PARAMS), (mean, 12), (scale, 12)):
            if a is None:
                arrays.append(None)
                continue
            arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32)).reshape(-1)
            if arr.size != size:
                raise ValueError(f"expected {size} values, got {arr.size}")
            arrays.append(arr)
        self._call(lib.gc_sched_set, *[None if a is None else a.ctypes.data_as(fp) for a in arrays])

    def predict(self, x) -> np.ndarray:
        x = _f32(x, (12,), "x")
        y = np.zeros((x.shape[0], 3), np.float32)
        fp = POINTER(c_float)
        self._call(lib.gc_sched_predict, x.ctypes.data_as(fp), x.shape[0], y.ctypes.data_as(fp))
        return y

    def loss_grad(self, x, t, mask) -> tuple[float, np.ndarray]:
        x, t, mask = _f32(x, (12,), "x"), _f32(t, (3,), "t"), _f32(mask, (3,), "mask")
        grad = np.zeros(SCHED_PARAMS, np.float32)
        loss = c_float()
        fp = POINTER(c_float)
        self._call(lib.gc_sched_loss_grad, x.ctypes.data_as(fp), t.ctypes.data_as(fp), mask.ctypes.data_as(fp),
                   x.shape[0], byref(loss), grad.ctypes.data_as(fp))
        return loss.value, grad

    def train(self, x, t, mask, steps: int = 1000, lr: float = 1e-2) -> float:
        x, t, mask = _f32(x, (12,), "x"), _f32(t, (3,), "t"), _f32(mask, (3,), "mask")
        loss = c_float()
        fp = POINTER(c_float)
        self._call(lib.gc_sched_train, x.ctypes.data_as(fp), t.ctypes.data_as(fp), mask.ctypes.data_as(fp),
                   x.shape[0], int(steps), c_float(lr), byref(loss))
        return loss.value

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            lib.gc_sched_free(h)
            self._handle = None
