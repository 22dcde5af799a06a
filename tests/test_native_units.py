"""Builds and runs the host-only C++ unit tests (module cache with a fake loader, launch
geometry, PageRank's binned layout replayed against direct row sums); no GPU involved."""

from __future__ import annotations

import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_native_units(tmp_path):
    exe = tmp_path / "units"
    src = os.path.join(REPO, "tests", "native", "test_units.cpp")
    inc = os.path.join(REPO, "grafeo-cuda_amd", "csrc", "runtime")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{inc}", src, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all native unit tests passed" in out.stdout


def test_pagerank_bin_layout(tmp_path):
    exe = tmp_path / "pr_bin"
    src = os.path.join(REPO, "tests", "native", "test_pr_bin.cpp")
    inc = os.path.join(REPO, "grafeo-cuda_amd", "csrc", "algos")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{inc}", src, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "pr_bin layout tests passed" in out.stdout
