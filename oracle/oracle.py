"""ctypes access to the CPU oracle (oracle.c) plus small pure-Python restatements.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, and there only as the checker. The product never imports it.
"""

from __future__ import annotations

import ctypes
import os
import re
import subprocess
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int32, c_size_t, c_uint32, c_uint64

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build() -> str:
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    return LIB_PATH


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    u32p, i32p = POINTER(c_uint32), POINTER(c_int32)
    sig = {
        "oracle_set_threads": (None, [c_int]),
        "oracle_max_threads": (c_int, []),
        "oracle_validate_csr": (c_int, [u32p, c_size_t, u32p, c_size_t, c_char_p, c_size_t]),
        "oracle_from_edge_pairs": (c_size_t, [u32p, c_size_t, c_uint32, c_int, u32p, u32p]),
        "oracle_from_edge_pairs_oriented": (c_size_t, [u32p, c_size_t, c_uint32, u32p, u32p]),
        "oracle_triangle_count_ref": (c_uint64, [u32p, u32p, c_uint32]),
        "oracle_triangle_count_oriented_ref": (c_uint64, [u32p, u32p, c_uint32]),
        "oracle_triangle_count_fast": (c_uint64, [u32p, u32p, c_uint32]),
        "oracle_bfs": (c_int, [u32p, u32p, c_uint32, c_uint32, i32p, u32p, i32p]),
        "oracle_pagerank_f64": (
            None,
            [u32p, u32p, c_uint32, c_double, c_uint32, c_double, POINTER(c_double), u32p, POINTER(c_int)],
        ),
        "oracle_k_truss_ref": (c_uint32, [u32p, u32p, c_uint32, c_uint32, POINTER(ctypes.c_uint8)]),
        "oracle_k_truss_fast": (c_uint32, [u32p, u32p, c_uint32, c_uint32, POINTER(ctypes.c_uint8)]),
        "oracle_truss_numbers": (ctypes.c_int64, [u32p, u32p, c_uint32, u32p]),
        "oracle_tc_prepare": (ctypes.c_void_p, [u32p, u32p, c_uint32]),
        "oracle_tc_count_rows": (c_uint64, [ctypes.c_void_p, c_uint32, c_uint32]),
        "oracle_tc_release": (None, [ctypes.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _p(a: np.ndarray, t=c_uint32):
    return a.ctypes.data_as(POINTER(t))


def set_threads(threads: int) -> None:
    lib.oracle_set_threads(int(threads))


def max_threads() -> int:
    return lib.oracle_max_threads()


def validate_csr(offsets, targets) -> str | None:
    """None if valid, else the InvalidCsr message."""
    off, tgt = _u32(offsets), _u32(targets)
    buf = ctypes.create_string_buffer(512)
    rc = lib.oracle_validate_csr(_p(off), off.size, _p(tgt), tgt.size, buf, 512)
    return None if rc == 0 else buf.value.decode()


def from_edge_pairs(edges, n: int, symmetric: bool):
    pairs = _u32(np.asarray(edges, dtype=np.int64).reshape(-1, 2)) if len(edges) else np.zeros((0, 2), np.uint32)
    off = np.zeros(n + 1, np.uint32)
    tgt = np.zeros(2 * len(pairs) + 1, np.uint32)
    m = lib.oracle_from_edge_pairs(_p(pairs), len(pairs), n, int(symmetric), _p(off), _p(tgt))
    return off, tgt[:m].copy()


def from_edge_pairs_oriented(edges, n: int):
    pairs = _u32(np.asarray(edges, dtype=np.int64).reshape(-1, 2)) if len(edges) else np.zeros((0, 2), np.uint32)
    off = np.zeros(n + 1, np.uint32)
    tgt = np.zeros(len(pairs) + 1, np.uint32)
    m = lib.oracle_from_edge_pairs_oriented(_p(pairs), len(pairs), n, _p(off), _p(tgt))
    return off, tgt[:m].copy()


def triangle_count(offsets, targets, fast: bool = False) -> int:
    off, tgt = _u32(offsets), _u32(targets)
    fn = lib.oracle_triangle_count_fast if fast else lib.oracle_triangle_count_ref
    return int(fn(_p(off), _p(tgt), off.size - 1))


class TriangleSample:
    """triangle_count(fast=True) in two timed halves (the CPU baseline's bounded sample): the
    degree-ordered DAG is built once; count_rows(stride, phase) merges only the DAG rows r with
    r % stride == phase, and the phases of one stride add up to the full count."""

    def __init__(self, offsets, targets):
        self._off, self._tgt = _u32(offsets), _u32(targets)
        self._h = lib.oracle_tc_prepare(_p(self._off), _p(self._tgt), self._off.size - 1)

    def count_rows(self, stride: int, phase: int = 0) -> int:
        return int(lib.oracle_tc_count_rows(self._h, int(stride), int(phase)))

    def close(self) -> None:
        if self._h:
            lib.oracle_tc_release(self._h)
            self._h = None

    def __del__(self):
        self.close()


def triangle_count_oriented(offsets, targets) -> int:
    off, tgt = _u32(offsets), _u32(targets)
    return int(lib.oracle_triangle_count_oriented_ref(_p(off), _p(tgt), off.size - 1))


def bfs(offsets, targets, source: int):
    """(distances int32[n], reachable, max_depth); ValueError for a source out of range."""
    off, tgt = _u32(offsets), _u32(targets)
    n = off.size - 1
    dist = np.zeros(max(n, 1), np.int32)
    reach, depth = c_uint32(), c_int32()
    rc = lib.oracle_bfs(_p(off), _p(tgt), n, source, _p(dist, c_int32), byref(reach), byref(depth))
    if rc != 0:
        raise ValueError(f"source node {source} out of range (num_nodes = {n})")
    return dist[:n].copy(), reach.value, depth.value


def pagerank(offsets, targets, damping=0.85, max_iterations=100, tolerance=1e-6):
    """(ranks float64[n], iterations, converged)."""
    off, tgt = _u32(offsets), _u32(targets)
    n = off.size - 1
    ranks = np.zeros(max(n, 1), np.float64)
    it, conv = c_uint32(), c_int()
    lib.oracle_pagerank_f64(_p(off), _p(tgt), n, damping, max_iterations, tolerance,
                            _p(ranks, c_double), byref(it), byref(conv))
    return ranks[:n].copy(), it.value, bool(conv.value)


def k_truss(offsets, targets, k: int, fast: bool = False):
    """(edges as list of (u, v) with u < v in CSR order, iterations)."""
    off, tgt = _u32(offsets), _u32(targets)
    n = off.size - 1
    m = int(off[-1])
    removed = np.zeros(max(m, 1), np.uint8)
    fn = lib.oracle_k_truss_fast if fast else lib.oracle_k_truss_ref
    iterations = int(fn(_p(off), _p(tgt), n, k, _p(removed, ctypes.c_uint8)))
    edges = []
    for u in range(n):
        for e in range(int(off[u]), int(off[u + 1])):
            v = int(tgt[e])
            if v > u and not removed[e]:
                edges.append((u, v))
    return edges, iterations


def truss_numbers(offsets, targets):
    """(edges as list of (u, v) with u < v in CSR order, trussness per edge as a uint32 array)
    of a symmetric CSR, by sequential peeling (oracle_truss_numbers)."""
    off, tgt = _u32(offsets), _u32(targets)
    n = off.size - 1
    truss = np.zeros(max(int(off[-1]), 1), np.uint32)
    count = int(lib.oracle_truss_numbers(_p(off), _p(tgt), n, _p(truss)))
    if count < 0:
        raise ValueError("truss numbers need a symmetric CSR")
    edges = [(u, int(v)) for u in range(n) for v in tgt[off[u]:off[u + 1]] if v > u]
    return edges, truss[:count].copy()


# Unicode's White_Space property, which Rust's str::trim and split_whitespace use (Python's
# str.isspace also counts U+001C..001F, so it is not used here).
RUST_WHITESPACE = ("\t\n\x0b\x0c\r \x85\xa0\u1680\u2000\u2001\u2002\u2003\u2004\u2005\u2006\u2007"
                   "\u2008\u2009\u200a\u2028\u2029\u202f\u205f\u3000")
_RUST_WS_RUN = re.compile("[" + re.escape(RUST_WHITESPACE) + "]+")


def parse_tsv_to_edges(path: str):
    """parse_tsv_to_edges (crates/grafeo-cuda-core/src/csr_export.rs:246-287) in Python:
    lines split at '\n' as BufRead::lines does (a line that is not UTF-8 fails the whole
    parse: UnicodeDecodeError here, an io::Error there); each trimmed of White_Space (str::trim);
    skip blank, '#' and '%' lines; the first two White_Space-separated fields as u64 (Rust's
    u64::from_str: optional '+', digits only, no overflow); ids remapped by first appearance;
    lines that do not parse are skipped."""

    def parse_u64(s: str):
        t = s[1:] if s.startswith("+") else s
        if not t or not t.isascii() or not t.isdigit():
            return None
        v = int(t)
        return v if v <= 0xFFFFFFFFFFFFFFFF else None

    ids: dict[int, int] = {}
    edges = []
    with open(path, "rb") as f:
        for raw in f.read().split(b"\n"):
            line = raw.decode("utf-8").strip(RUST_WHITESPACE)
            if not line or line.startswith("#") or line.startswith("%"):
                continue
            parts = _RUST_WS_RUN.split(line)
            if len(parts) < 2:
                continue
            s, d = parse_u64(parts[0]), parse_u64(parts[1])
            if s is None or d is None:
                continue
            si = ids.setdefault(s, len(ids))
            di = ids.setdefault(d, len(ids))
            edges.append((si, di))
    return edges, len(ids)
