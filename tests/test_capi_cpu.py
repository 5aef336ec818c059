"""The C-ABI library loads, exports every entry include/gpusort_hip.h declares,
and reports errors through status codes and gs_last_error (no GPU needed:
no compute call is made)."""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

import gpusort_amd as g

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "gpusort_hip.h"


def declared_entries() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(gs_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_entry():
    lib = g.load_library()
    names = declared_entries()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(g.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (gs_\w+)", out))
    assert set(names) <= exported
    # the Python binding knows every entry and nothing undeclared
    assert set(g.exported_symbols()) == set(names)


def test_library_is_gfx950_code():
    # The embedded device code object names its target.
    blob = g.LIB_PATH.read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_version_and_error_channel():
    lib = g.load_library()
    assert b"gfx950" in lib.gs_version()
    st = lib.gs_sort_keys(None, g.U32, None, 10, 0)
    assert st == 1
    assert b"null handle" in lib.gs_last_error()


def test_no_device_paths_on_cpu(gpu_available):
    if gpu_available:
        pytest.skip("a device is visible")
    assert g.device_count() == 0
    lib = g.load_library()
    h = ctypes.c_void_p()
    st = lib.gs_create(ctypes.byref(h), 0)
    assert st == 5  # GS_NO_DEVICE
    assert b"No HIP-capable devices were detected" in lib.gs_last_error()
    with pytest.raises(g.GpusortError) as e:
        g.Handle(0)
    assert e.value.status == 5


def test_argument_validation_without_device():
    lib = g.load_library()
    assert lib.gs_get_device_count(None) == 1
    assert lib.gs_destroy(None) == 0
    assert lib.gs_vrank_sort(None, 0, g.U32, None, g.NONE, None, None, 0, 8, 32, 0, None) == 1
    assert lib.gs_gather(None, 4, None, None, None, 0) == 1
    assert lib.gs_profile_end(None, None, 0, None) == 1


def test_server_needs_a_device(gpu_available):
    if gpu_available:
        pytest.skip("a device is visible")
    with pytest.raises(g.GpusortError) as e:
        g.Server()
    assert e.value.status == 5
    lib = g.load_library()
    assert lib.gs_server_submit(None, g.U32, None, g.NONE, None, 0, 0, None) == 1
    assert lib.gs_server_destroy(None) == 0
