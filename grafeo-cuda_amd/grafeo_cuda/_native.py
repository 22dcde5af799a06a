"""ctypes binding of libgrafeo_hip.so (declared in include/grafeo_hip.h).

Low-level layer under the `grafeo_cuda` classes: loads the in-tree library, declares every
entry point's signature, turns non-zero gc_status codes into exceptions and converts arrays.
Importing fails loudly when the library is not built; there is no fallback implementation.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_float, c_int, c_int32, c_size_t
from ctypes import c_uint32, c_uint64, c_void_p

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgrafeo_hip.so")

# gc_status values (include/grafeo_hip.h); names follow CudaError's variants
# (crates/grafeo-cuda-core/src/error.rs:5-39).
STATUS_NAMES = {
    0: "Ok",
    1: "NoDevice",
    2: "Driver",
    3: "OutOfMemory",
    4: "Kernel",
    5: "Export",
    6: "InvalidCsr",
}


class GcError(RuntimeError):
    """A failed library call; `status` is the gc_status code, str() the CudaError text."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status
        self.kind = STATUS_NAMES.get(status, str(status))


class Csr(ctypes.Structure):
    _fields_ = [
        ("offsets", POINTER(c_uint32)),
        ("targets", POINTER(c_uint32)),
        ("num_nodes", c_uint32),
        ("num_edges", c_uint32),
    ]


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python grafeo-cuda_amd/build.py` "
            "(or __graft_entry__.build()) first"
        )
    lib = ctypes.CDLL(LIB_PATH)
    u32p, i32p, f32p = POINTER(c_uint32), POINTER(c_int32), POINTER(c_float)
    sig = {
        "gc_last_error": (c_char_p, []),
        "gc_free": (None, [c_void_p]),
        "gc_device_count": (c_int, [u32p]),
        "gc_device_new": (c_int, [c_uint32, POINTER(c_void_p)]),
        "gc_device_retain": (c_void_p, [c_void_p]),
        "gc_device_release": (None, [c_void_p]),
        "gc_device_properties": (c_int, [c_void_p, c_char_p, c_size_t, POINTER(c_uint64), u32p]),
        "gc_csr_free": (None, [POINTER(Csr)]),
        "gc_validate_csr": (c_int, [u32p, c_size_t, u32p, c_size_t]),
        "gc_csr_validate": (c_int, [POINTER(Csr)]),
        "gc_csr_from_edge_pairs": (c_int, [u32p, c_size_t, c_uint32, c_int, POINTER(Csr)]),
        "gc_csr_from_edge_pairs_device": (c_int, [c_void_p, u32p, c_size_t, c_uint32, c_int, POINTER(Csr)]),
        "gc_csr_from_edge_pairs_oriented_device": (c_int, [c_void_p, u32p, c_size_t, c_uint32, POINTER(Csr)]),
        "gc_parse_tsv_to_edges_device": (c_int, [c_void_p, c_char_p, POINTER(POINTER(c_uint32)), POINTER(c_size_t),
                                                 POINTER(c_uint32)]),
        "gc_graph_from_tsv": (c_int, [c_void_p, c_char_p, c_int32, POINTER(c_void_p), POINTER(c_size_t)]),
        "gc_graph_from_edge_pairs": (c_int, [c_void_p, u32p, c_size_t, c_uint32, c_int, POINTER(c_void_p)]),
        "gc_graph_relabelled": (c_int, [c_void_p, POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p),
                                        POINTER(c_void_p)]),
        "gc_csr_from_edge_pairs_oriented": (c_int, [u32p, c_size_t, c_uint32, POINTER(Csr)]),
        "gc_parse_tsv_to_edges": (c_int, [c_char_p, POINTER(u32p), POINTER(c_size_t), u32p]),
        "gc_graph_from_csr": (c_int, [c_void_p, u32p, c_size_t, u32p, c_size_t, POINTER(c_void_p)]),
        "gc_graph_from_csr_arrays": (c_int, [c_void_p, POINTER(Csr), POINTER(c_void_p)]),
        "gc_graph_free": (None, [c_void_p]),
        "gc_graph_num_nodes": (c_uint32, [c_void_p]),
        "gc_graph_num_edges": (c_uint32, [c_void_p]),
        "gc_graph_device_offsets": (c_void_p, [c_void_p]),
        "gc_graph_device_targets": (c_void_p, [c_void_p]),
        "gc_graph_prepare": (c_int, [c_void_p, POINTER(c_double)]),
        "gc_graph_is_symmetric": (c_int, [c_void_p, i32p]),
        "gc_graph_dag_stats": (c_int, [c_void_p, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]),
        "gc_device_malloc": (c_int, [c_void_p, c_size_t, POINTER(c_void_p)]),
        "gc_device_free": (c_int, [c_void_p, c_void_p]),
        "gc_device_copy_to_host": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
        "gc_triangle_count": (c_int, [c_void_p, POINTER(c_uint64), POINTER(c_double)]),
        "gc_triangle_count_noalloc": (c_int, [c_void_p, POINTER(c_uint64)]),
        "gc_triangle_count_oriented": (c_int, [c_void_p, POINTER(c_uint64), POINTER(c_double)]),
        "gc_bfs": (c_int, [c_void_p, c_uint32, i32p, u32p, i32p, POINTER(c_double)]),
        "gc_bfs_device": (c_int, [c_void_p, c_uint32, c_void_p, u32p, i32p, POINTER(c_double)]),
        "gc_pagerank": (
            c_int,
            [c_void_p, c_float, c_uint32, c_float, f32p, u32p, i32p, POINTER(c_double)],
        ),
        "gc_pagerank_device": (
            c_int,
            [c_void_p, c_float, c_uint32, c_float, c_void_p, u32p, i32p, POINTER(c_double)],
        ),
        "gc_pagerank_layout": (c_int, [c_void_p, POINTER(c_uint64), c_uint32]),
        "gc_k_truss": (
            c_int,
            [c_void_p, c_uint32, POINTER(u32p), POINTER(c_size_t), u32p, POINTER(c_double)],
        ),
        "gc_msbfs": (
            c_int,
            [c_void_p, u32p, c_uint32, i32p, u32p, i32p, POINTER(c_double)],
        ),
        "gc_msbfs_device": (
            c_int,
            [c_void_p, u32p, c_uint32, c_void_p, u32p, i32p, POINTER(c_double)],
        ),
        "gc_batch_from_csrs": (
            c_int,
            [c_void_p, c_uint32, POINTER(u32p), u32p, POINTER(u32p), POINTER(c_uint64), POINTER(c_void_p)],
        ),
        "gc_batch_free": (None, [c_void_p]),
        "gc_batch_num_graphs": (c_uint32, [c_void_p]),
        "gc_batch_triangle_count": (c_int, [c_void_p, POINTER(c_uint64), POINTER(c_double)]),
        "gc_batch_set_config": (c_int, [c_void_p, u32p]),
        "gc_batch_get_config": (c_int, [c_void_p, u32p]),
        "gc_batch_tune": (c_int, [c_void_p, u32p, POINTER(c_double)]),
        "gc_sched_new": (c_int, [c_void_p, c_uint64, POINTER(c_void_p)]),
        "gc_sched_free": (None, [c_void_p]),
        "gc_sched_get": (c_int, [c_void_p, f32p, f32p, f32p]),
        "gc_sched_set": (c_int, [c_void_p, f32p, f32p, f32p]),
        "gc_sched_predict": (c_int, [c_void_p, f32p, c_uint32, f32p]),
        "gc_sched_loss_grad": (c_int, [c_void_p, f32p, f32p, f32p, c_uint32, f32p, f32p]),
        "gc_sched_train": (c_int, [c_void_p, f32p, f32p, f32p, c_uint32, c_uint32, c_float, f32p]),
        "gc_graph_set_scheduler": (c_int, [c_void_p, c_void_p, c_void_p]),
        "gc_graph_last_schedule": (c_int, [c_void_p, u32p, c_uint32, u32p]),
        "gc_bfs_level_timings": (c_int, [c_void_p, c_uint32, c_uint32, f32p, f32p, u32p]),
        "gc_tc_bucket_timings": (c_int, [c_void_p, c_uint32, f32p, f32p, u32p]),
        "gc_truss_numbers": (
            c_int,
            [c_void_p, POINTER(u32p), POINTER(u32p), POINTER(c_size_t), u32p, POINTER(c_double)],
        ),
        "gc_profile_enable": (c_int, [c_void_p, c_int32]),
        "gc_profile_query": (c_int, [c_void_p, c_char_p, POINTER(c_double), POINTER(c_uint64)]),
        "gc_profile_only": (c_int, [c_void_p, c_char_p]),
        "gc_gen_erdos_renyi": (c_int, [c_uint32, c_uint64, c_uint64, POINTER(u32p), POINTER(c_size_t)]),
        "gc_gen_rmat": (
            c_int,
            [c_uint32, c_uint32, c_double, c_double, c_double, c_uint64, c_int32, POINTER(u32p),
             POINTER(c_size_t)],
        ),
        "gc_gen_grid": (c_int, [c_uint32, c_uint32, POINTER(u32p), POINTER(c_size_t)]),
        "gc_code_object_count": (c_uint32, []),
        "gc_code_object_info": (
            c_int,
            [c_uint32, POINTER(c_char_p), POINTER(c_char_p), POINTER(c_char_p), POINTER(c_uint64)],
        ),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# Entry points declared in include/grafeo_hip.h (checked by the tests against the header).
EXPORTED = [
    "gc_last_error", "gc_free", "gc_device_count", "gc_device_new", "gc_device_retain",
    "gc_device_release", "gc_device_properties", "gc_csr_free", "gc_validate_csr",
    "gc_csr_validate", "gc_csr_from_edge_pairs", "gc_csr_from_edge_pairs_oriented",
    "gc_csr_from_edge_pairs_device", "gc_csr_from_edge_pairs_oriented_device", "gc_parse_tsv_to_edges_device",
    "gc_graph_from_tsv", "gc_graph_from_edge_pairs",
    "gc_graph_relabelled",
    "gc_parse_tsv_to_edges", "gc_graph_from_csr", "gc_graph_from_csr_arrays", "gc_graph_free",
    "gc_graph_num_nodes", "gc_graph_num_edges", "gc_graph_device_offsets",
    "gc_graph_device_targets", "gc_graph_prepare", "gc_graph_is_symmetric", "gc_graph_dag_stats",
    "gc_device_malloc", "gc_device_free", "gc_device_copy_to_host", "gc_triangle_count", "gc_triangle_count_noalloc",
    "gc_triangle_count_oriented", "gc_bfs", "gc_bfs_device", "gc_pagerank", "gc_pagerank_device", "gc_pagerank_layout",
    "gc_k_truss", "gc_truss_numbers", "gc_msbfs", "gc_msbfs_device", "gc_batch_from_csrs",
    "gc_batch_free", "gc_batch_num_graphs", "gc_batch_triangle_count", "gc_batch_set_config",
    "gc_batch_get_config", "gc_batch_tune", "gc_sched_new", "gc_sched_free", "gc_sched_get",
    "gc_sched_set", "gc_sched_predict", "gc_sched_loss_grad", "gc_sched_train",
    "gc_graph_set_scheduler", "gc_graph_last_schedule", "gc_bfs_level_timings",
    "gc_tc_bucket_timings", "gc_profile_enable", "gc_profile_query", "gc_profile_only", "gc_gen_erdos_renyi", "gc_gen_rmat",
    "gc_gen_grid", "gc_code_object_count", "gc_code_object_info",
]


def check(status: int) -> None:
    if status != 0:
        raise GcError(status, lib.gc_last_error().decode(errors="replace"))


def as_u32(values, what: str) -> np.ndarray:
    """Contiguous uint32 copy-free view when possible; rejects values outside u32."""
    if isinstance(values, np.ndarray):
        if values.dtype == np.uint32:
            return np.ascontiguousarray(values)
        if values.dtype.kind not in "iu":
            raise TypeError(f"{what} must hold integers, got dtype {values.dtype}")
        arr = values
    else:
        try:
            arr = np.asarray(list(values), dtype=object)
        except TypeError as e:
            raise TypeError(f"{what} must be a sequence of integers") from e
        if arr.size and not all(isinstance(v, (int, np.integer)) and not isinstance(v, bool) for v in arr):
            raise TypeError(f"{what} must be a sequence of integers")
        arr = np.asarray([int(v) for v in arr], dtype=np.int64) if arr.size else np.zeros(0, np.int64)
    if arr.size and (int(arr.min()) < 0 or int(arr.max()) > 0xFFFFFFFF):
        raise OverflowError(f"{what} holds a value outside the u32 range")
    return np.ascontiguousarray(arr, dtype=np.uint32)


def u32_ptr(arr: np.ndarray):
    return arr.ctypes.data_as(POINTER(c_uint32))


def take_u32(ptr, count: int) -> np.ndarray:
    """Copies `count` u32 from a library-allocated array and frees it."""
    try:
        if count == 0:
            return np.zeros(0, dtype=np.uint32)
        return np.ctypeslib.as_array(ptr, shape=(count,)).copy()
    finally:
        lib.gc_free(ctypes.cast(ptr, c_void_p))


def csr_to_numpy(csr: Csr) -> tuple[np.ndarray, np.ndarray]:
    try:
        n, m = csr.num_nodes, csr.num_edges
        offsets = np.ctypeslib.as_array(csr.offsets, shape=(n + 1,)).copy()
        targets = (
            np.ctypeslib.as_array(csr.targets, shape=(m,)).copy() if m else np.zeros(0, np.uint32)
        )
        return offsets, targets
    finally:
        lib.gc_csr_free(byref(csr))


def pairs_array(edges) -> np.ndarray:
    """(E, 2) uint32 array from a sequence of pairs or an array."""
    if isinstance(edges, np.ndarray):
        arr = edges.reshape(-1, 2) if edges.size else np.zeros((0, 2), np.int64)
    else:
        edges = list(edges)
        arr = np.asarray(edges, dtype=np.int64).reshape(-1, 2) if edges else np.zeros((0, 2), np.int64)
    return as_u32(arr.reshape(-1), "edges").reshape(-1, 2)


# ---- CSR model (crates/grafeo-cuda-core/src/csr_export.rs) -------------------


def validate_csr(offsets, targets) -> None:
    off, tgt = as_u32(offsets, "offsets"), as_u32(targets, "targets")
    check(lib.gc_validate_csr(u32_ptr(off), off.size, u32_ptr(tgt), tgt.size))


def from_edge_pairs(edges, num_nodes: int, symmetric: bool):
    pairs = pairs_array(edges)
    csr = Csr()
    check(lib.gc_csr_from_edge_pairs(u32_ptr(pairs), pairs.shape[0], num_nodes, int(bool(symmetric)), byref(csr)))
    return csr_to_numpy(csr)


def from_edge_pairs_device(device_handle, edges, num_nodes: int, symmetric: bool):
    """from_edge_pairs built on the GPU of `device_handle` (same arrays)."""
    pairs = pairs_array(edges)
    csr = Csr()
    check(lib.gc_csr_from_edge_pairs_device(device_handle, u32_ptr(pairs), pairs.shape[0], num_nodes,
                                            int(bool(symmetric)), byref(csr)))
    return csr_to_numpy(csr)


def from_edge_pairs_oriented_device(device_handle, edges, num_nodes: int):
    """from_edge_pairs_oriented built on the GPU of `device_handle` (same arrays)."""
    pairs = pairs_array(edges)
    csr = Csr()
    check(lib.gc_csr_from_edge_pairs_oriented_device(device_handle, u32_ptr(pairs), pairs.shape[0], num_nodes,
                                                     byref(csr)))
    return csr_to_numpy(csr)


def from_edge_pairs_oriented(edges, num_nodes: int):
    pairs = pairs_array(edges)
    csr = Csr()
    check(lib.gc_csr_from_edge_pairs_oriented(u32_ptr(pairs), pairs.shape[0], num_nodes, byref(csr)))
    return csr_to_numpy(csr)


def parse_tsv_to_edges_device(device_handle, path: str):
    """parse_tsv_to_edges on the GPU of `device_handle` (same pairs and node count)."""
    ptr, count, nodes = POINTER(c_uint32)(), c_size_t(), c_uint32()
    check(lib.gc_parse_tsv_to_edges_device(device_handle, os.fsencode(path), byref(ptr), byref(count), byref(nodes)))
    return take_u32(ptr, 2 * count.value).reshape(-1, 2), nodes.value


def parse_tsv_to_edges(path: str):
    ptr, count, nodes = POINTER(c_uint32)(), c_size_t(), c_uint32()
    check(lib.gc_parse_tsv_to_edges(os.fsencode(path), byref(ptr), byref(count), byref(nodes)))
    return take_u32(ptr, 2 * count.value).reshape(-1, 2), nodes.value


# ---- generators ---------------------------------------------------------------


def gen_erdos_renyi(n: int, num_draws: int, seed: int) -> np.ndarray:
    ptr, count = POINTER(c_uint32)(), c_size_t()
    check(lib.gc_gen_erdos_renyi(n, num_draws, seed, byref(ptr), byref(count)))
    return take_u32(ptr, 2 * count.value).reshape(-1, 2)


def gen_rmat(scale: int, edge_factor: int, seed: int, a=0.57, b=0.19, c=0.19, permute=True) -> np.ndarray:
    ptr, count = POINTER(c_uint32)(), c_size_t()
    check(lib.gc_gen_rmat(scale, edge_factor, a, b, c, seed, int(bool(permute)), byref(ptr), byref(count)))
    return take_u32(ptr, 2 * count.value).reshape(-1, 2)


def gen_grid(rows: int, cols: int) -> np.ndarray:
    ptr, count = POINTER(c_uint32)(), c_size_t()
    check(lib.gc_gen_grid(rows, cols, byref(ptr), byref(count)))
    return take_u32(ptr, 2 * count.value).reshape(-1, 2)


def code_objects() -> list[dict]:
    out = []
    for i in range(lib.gc_code_object_count()):
        name, src, obj, size = c_char_p(), c_char_p(), c_char_p(), c_uint64()
        check(lib.gc_code_object_info(i, byref(name), byref(src), byref(obj), byref(size)))
        out.append(
            {
                "name": name.value.decode(),
                "source_sha256": src.value.decode(),
                "object_sha256": obj.value.decode(),
                "size": size.value,
            }
        )
    return out


def device_count() -> int:
    c = c_uint32()
    check(lib.gc_device_count(byref(c)))
    return c.value
