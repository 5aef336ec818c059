"""Shared fixtures.  GPU tests are marked ``gpu`` and need a visible device."""
from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def _has_gpu() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_available() -> bool:
    return _has_gpu()


@pytest.fixture(scope="session")
def handle():
    if not _has_gpu():
        pytest.skip("no HIP device")
    import gpusort_amd

    h = gpusort_amd.Handle(0)
    yield h
    h.close()


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no HIP device")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
