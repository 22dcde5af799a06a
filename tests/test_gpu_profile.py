"""Per-kernel event timing (gc_profile_enable, gc_profile_only, gc_profile_query): every launch is
timed by default, only the named kernels when a list is given, and the totals count each timed
launch once."""

from __future__ import annotations

import pytest

import grafeo_cuda
from graphs import rmat_graph

pytestmark = pytest.mark.gpu


def test_event_timing_of_all_and_of_named_kernels(device):
    off, tgt = rmat_graph(12, 16, seed=1)
    g = grafeo_cuda.GpuGraph.from_csr(off, tgt, device)
    g.prepare()
    expected = g.triangle_count()["total_triangles"]

    device.profile(True)
    assert g.triangle_count()["total_triangles"] == expected
    ms, launches = device.kernel_time("tc_bitmap16")
    assert launches == 1 and ms > 0.0
    assert device.kernel_time("fill_u32")[1] >= 1  # the work counters' reset is timed too

    device.profile(True, only=["tc_bitmap16"])
    for _ in range(3):
        assert g.triangle_count()["total_triangles"] == expected
    assert device.kernel_time("tc_bitmap16")[1] == 3
    assert device.kernel_time("fill_u32") == (0.0, 0)

    device.profile(True)  # the filter is cleared: every kernel again
    g.triangle_count()
    assert device.kernel_time("fill_u32")[1] >= 1
    device.profile(False)
    g.triangle_count()
    assert device.kernel_time("tc_bitmap16") == (0.0, 0)
