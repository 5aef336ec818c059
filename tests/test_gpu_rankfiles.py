"""The rank-file harness on the GPU: the reference's known-answer sequences
written in its on-disk format, sorted across virtual ranks and checked from
the output files: a shuffled arithmetic progression comes back as a1 + i*d
exactly (src/test/arithmetic_progression_checker.h:49-106), co-shuffled value
vectors that are functions of the keys stay attached to their keys
(src/test/test_key_value.cc:599-683), and every output passes CheckValid."""
from __future__ import annotations

import numpy as np
import pytest

from gpusort_amd import rankfiles as rf
from gpu_util import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("algo,k", [("hyksort", 4), ("samplesort", 0), ("hyperquicksort", 0)])
@pytest.mark.parametrize("descending", [False, True], ids=["asc", "desc"])
def test_arithmetic_progression_files(tmp_path, algo, k, descending):
    a1, d, n, np_ = -1000, 7, 50_000, 8
    ap = a1 + d * np.arange(n, dtype=np.int64)
    shuffled = ap[np.random.default_rng(3).permutation(n)]
    src, out = tmp_path / "in", tmp_path / "out"
    src.mkdir()
    rf.save_data(np_, str(src), ".dat", shuffled)
    say(f"rank files AP algo={algo} desc={descending}")
    assert rf.sort_rank_files(str(src), np_, "int64", [], algo, max(k, 2), 32, descending, str(out))
    got, off = rf.read_ranks(np_, str(out), ".dat", np.int64)
    assert np.array_equal(got, ap[::-1] if descending else ap)
    assert (np.diff(off) > 0).all()


def test_key_value_files_values_follow_keys(tmp_path):
    n, np_ = 120_000, 16
    keys = (np.random.default_rng(4).integers(0, 30_000, n)).astype(np.uint32)
    v1 = keys.astype(np.float64) * 2.0 + 0.5
    v2 = -keys.astype(np.int32)
    src, out = tmp_path / "in", tmp_path / "out"
    src.mkdir()
    rf.save_data(np_, str(src), ".key", keys)
    rf.save_data(np_, str(src), ".v1", v1)
    rf.save_data(np_, str(src), ".v2", v2)
    say("rank files KV two value vectors")
    assert rf.sort_rank_files(str(src), np_, "uint32", ["float64", "int32"], "hyksort", 4, 32, False, str(out))
    gk, _ = rf.read_ranks(np_, str(out), ".key", np.uint32)
    g1, _ = rf.read_ranks(np_, str(out), ".v1", np.float64)
    g2, _ = rf.read_ranks(np_, str(out), ".v2", np.int32)
    assert np.array_equal(gk, np.sort(keys))
    assert np.array_equal(g1, gk.astype(np.float64) * 2.0 + 0.5)
    assert np.array_equal(g2, -gk.astype(np.int32))


def test_single_value_file_is_lexicographic(tmp_path):
    n, np_ = 40_000, 4
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 500, n).astype(np.uint64)
    vals = rng.integers(0, 1 << 31, n).astype(np.uint32)
    src, out = tmp_path / "in", tmp_path / "out"
    src.mkdir()
    rf.save_data(np_, str(src), ".key", keys)
    rf.save_data(np_, str(src), ".v1", vals)
    say("rank files KV one value vector")
    assert rf.sort_rank_files(str(src), np_, "uint64", ["uint32"], "samplesort", 2, 32, False, str(out))
    gk, _ = rf.read_ranks(np_, str(out), ".key", np.uint64)
    gv, _ = rf.read_ranks(np_, str(out), ".v1", np.uint32)
    order = np.lexsort((vals, keys))
    assert np.array_equal(gk, keys[order]) and np.array_equal(gv, vals[order])
