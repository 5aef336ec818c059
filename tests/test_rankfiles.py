"""Per-rank data files in the reference's on-disk format (int64 count + raw
elements at <dir>/<rank><ext>): the Python reader/writer, the C++ header
include/gpusort/io/rank_files.h, and each reading what the other wrote."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from gpusort_amd import rankfiles as rf

PROG = Path(__file__).resolve().parents[1] / "tests" / "cpp" / "dropin_test"


def test_block_rule_known_answers():
    # include/common.h:146-154
    assert rf.calc_num_elems_each_block(10, 4) == 3
    assert rf.calc_num_elems_each_block(9, 4) == 2   # tail element left unwritten by SaveData
    assert rf.calc_num_elems_each_block(12, 4) == 3
    assert rf.calc_num_elems_each_block(7, 3) == 3
    assert rf.calc_num_elems_each_block(1, 1) == 1


def test_byte_layout_and_round_trip(tmp_path):
    a = np.arange(10, dtype=np.uint32) * 3
    sizes = rf.save_data(4, str(tmp_path), ".dat", a)
    assert sizes == [3, 3, 3, 1]
    raw = (tmp_path / "0.dat").read_bytes()
    assert raw[:8] == np.int64(3).tobytes() and raw[8:] == a[:3].tobytes()
    back, off = rf.read_ranks(4, str(tmp_path), ".dat", np.uint32)
    assert np.array_equal(back, a) and off.tolist() == [0, 3, 6, 9, 10]
    assert rf.save_data(4, str(tmp_path), ".nine", np.zeros(9, np.int64)) == [2, 2, 2, 2]


def test_truncated_file_rejected(tmp_path):
    (tmp_path / "0.dat").write_bytes(np.int64(5).tobytes() + np.zeros(3, np.uint32).tobytes())
    with pytest.raises(ValueError):
        rf.read_data(0, str(tmp_path), ".dat", np.uint32)


def test_check_valid():
    assert rf.check_valid([np.array([1, 2]), np.array([2, 5])], True)
    assert not rf.check_valid([np.array([1, 3]), np.array([2, 5])], True)
    assert not rf.check_valid([np.array([1, 2]), np.array([], dtype=np.int64)], True)
    assert rf.check_valid([np.array([5, 4]), np.array([4, 1])], False)


def test_cpp_header_writes_what_python_reads(tmp_path):
    if not PROG.exists():
        pytest.fail(f"{PROG} is not built (run make -C tests/cpp)")
    r = subprocess.run([str(PROG), "rankfiles", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 4
    back, off = rf.read_ranks(4, str(tmp_path), ".dat", np.uint32)
    assert np.array_equal(back, np.arange(100, 110, dtype=np.uint32)) and off.tolist() == [0, 3, 6, 9, 10]
    v, voff = rf.read_ranks(3, str(tmp_path), ".v1", np.float64)
    assert v.tolist() == [0.5, -1.0, 2.25, 8.0, 3.5] and voff.tolist() == [0, 2, 2, 5]
    assert sorted(os.listdir(tmp_path))[:4] == ["0.dat", "0.nine", "0.v1", "1.dat"]
