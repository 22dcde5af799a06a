"""Shared test setup: import paths, the `gpu` marker and the device fixture."""

from __future__ import annotations

import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("grafeo-cuda_amd", "oracle", "tests"):
    path = os.path.join(REPO, sub)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")
    config.addinivalue_line("markers", "slow: large CPU-side checks")


@pytest.fixture(scope="session")
def device():
    import grafeo_cuda

    return grafeo_cuda.GpuDevice(0)
