"""The allocation-free triangle count (gc_triangle_count_noalloc) against gc_triangle_count, the
entry point it is a variant of, through the C ABI: bit-exact counts on two calls in a row, the
second with a smaller graph, before and after the 16-bit lists of the default kernel exist; and
its argument checks (unprepared or directed graph, null output), which fail before any device
work."""

from __future__ import annotations

from ctypes import byref, c_double, c_uint64

import pytest

import grafeo_cuda
import oracle
from grafeo_cuda import _native
from grafeo_cuda._native import GcError, check, lib
from graphs import er_graph, rmat_graph

pytestmark = pytest.mark.gpu


def count(g) -> int:
    total, secs = c_uint64(), c_double()
    check(lib.gc_triangle_count(g._handle, byref(total), byref(secs)))
    return total.value


def count_noalloc(g) -> int:
    total = c_uint64(12345)
    check(lib.gc_triangle_count_noalloc(g._handle, byref(total)))
    return total.value


def prepared(device, off, tgt):
    g = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
    g.prepare()
    return g


def test_noalloc_count_matches_triangle_count_on_two_calls(device):
    big_off, big_tgt = rmat_graph(14, 16, 3)
    small_off, small_tgt = er_graph(700, 8, 5)
    big = prepared(device, big_off, big_tgt)
    small = prepared(device, small_off, small_tgt)
    # first calls: the 16-bit lists are not built yet, so the 32-bit bitmap kernel runs
    first_big = count_noalloc(big)
    first_small = count_noalloc(small)
    expected_big = count(big)  # builds the 16-bit lists of `big`
    expected_small = count(small)
    print(f"rmat14: {expected_big} triangles, er700: {expected_small}", flush=True)
    assert first_big == expected_big
    assert first_small == expected_small
    assert expected_big == oracle.triangle_count(big_off, big_tgt)
    assert expected_small == oracle.triangle_count(small_off, small_tgt)
    # again, with the lists in place (tc_bitmap16), larger graph then smaller graph
    assert count_noalloc(big) == expected_big
    assert count_noalloc(small) == expected_small
    assert count_noalloc(small) == expected_small


def test_noalloc_count_on_rmat16(device):
    off, tgt = rmat_graph(16, 16, 7)
    g = prepared(device, off, tgt)
    expected = count(g)
    assert count_noalloc(g) == expected
    assert expected == oracle.triangle_count(off, tgt)


def test_noalloc_count_argument_checks(device):
    off, tgt = er_graph(300, 6, 9)
    g = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
    with pytest.raises(GcError) as e:
        count_noalloc(g)
    assert e.value.kind == "Kernel" and "gc_graph_prepare" in str(e.value)
    g.prepare()
    with pytest.raises(GcError) as e:
        check(lib.gc_triangle_count_noalloc(g._handle, None))
    assert e.value.kind == "Kernel"
    assert count_noalloc(g) == count(g)
    # a directed CSR: 0 -> 1 -> 2 -> 0
    d = grafeo_cuda.GpuGraph.from_csr([0, 1, 2, 3], [1, 2, 0], device)
    d.prepare()
    with pytest.raises(GcError) as e:
        count_noalloc(d)
    assert e.value.kind == "Kernel" and "symmetric" in str(e.value)
    # a graph without edges gives 0, prepared or not
    empty = grafeo_cuda.GpuGraph.from_csr([0, 0, 0], [], device)
    assert count_noalloc(empty) == 0
    assert _native.STATUS_NAMES[0] == "Ok"
