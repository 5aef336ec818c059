"""The C++ drop-in headers (include/gpusort/) compile against the reference's
API names and behave like it: built by tests/cpp/Makefile, the program checks
every HykSort / SampleSort / HyperQuickSort / CudaUtils / ParUtils overload it
calls against std::sort / std::stable_sort on the host."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PROG = ROOT / "tests" / "cpp" / "dropin_test"


def _program() -> Path:
    if not PROG.exists():
        pytest.fail(f"{PROG} is not built (run make -C tests/cpp)")
    return PROG


def test_dropin_without_device_throws_reference_exceptions(gpu_available):
    if gpu_available:
        pytest.skip("a device is visible")
    r = subprocess.run([str(_program()), "nodevice"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 4


@pytest.mark.gpu
def test_dropin_cases_on_device():
    r = subprocess.run([str(_program())], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 26
