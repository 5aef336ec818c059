"""Multisplit (gs_multisplit): stable bucket partition of unsorted keys by the
splitters' bucket table, against the oracle's restatement (std::lower_bound /
std::upper_bound bucket + stable counting partition), and its offsets against
gs_bucket_bounds of the sorted keys."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from gpusort_amd import datagen
from gpu_util import assert_bits_equal, dev, host, say

pytestmark = pytest.mark.gpu


def keys_of(dtype, n, seed, distinct=None):
    d = np.dtype(dtype)
    r = datagen.draws(seed, n)
    if distinct:
        r = datagen.draws(seed + 1, distinct)[r % np.uint64(distinct)]
    if d.kind == "f":
        return ((r >> np.uint64(40)).astype(np.float64) - 2.0**23).astype(d)
    if d.itemsize == 4:
        return (r >> np.uint64(32)).astype(np.uint32).view(d)
    return r.view(d)


def splitters_from(keys, nspl, seed, descending):
    if nspl == 0 or keys.size == 0:
        return keys[:0].copy()
    pick = (datagen.draws(seed, nspl) % np.uint64(keys.size)).astype(np.int64)
    s = np.sort(keys[pick])  # ties among splitters and with keys on purpose
    return s[::-1].copy() if descending else s


def run(handle, kdt, vdt, n, nspl, upper, descending, seed, distinct=None):
    keys = keys_of(kdt, n, seed, distinct)
    spl = splitters_from(keys, nspl, seed + 7, descending)
    say(f"multisplit kdt={np.dtype(kdt).name} vdt={vdt and np.dtype(vdt).name} n={n} nspl={nspl} "
        f"upper={upper} desc={descending} seed={seed}")
    check(handle, keys, vdt, spl, upper, descending)


def check(handle, keys, vdt, spl, upper, descending):
    n = keys.size
    vals = None if vdt is None else np.arange(n, dtype=np.uint64).astype(vdt)
    tk = dev(keys)
    ok = torch.empty_like(tk)
    tv = None if vals is None else dev(vals)
    ov = None if vals is None else torch.empty_like(tv)
    offs = handle.multisplit(tk, spl, ok, tv, ov, upper=upper, descending=descending)
    idx, want_offs = oracle.multisplit(keys, spl, upper, descending)
    assert np.array_equal(offs.astype(np.int64), want_offs), (offs[:10], want_offs[:10])
    assert_bits_equal(host(ok), keys[idx], "keys")
    if vals is not None:
        assert_bits_equal(host(ov), vals[idx], "payloads")
    assert_bits_equal(host(tk), keys, "input untouched")


@pytest.mark.parametrize("kdt,vdt", [(np.uint32, None), (np.uint32, np.uint32), (np.int32, np.uint64),
                                     (np.float32, np.uint32), (np.uint64, np.uint32), (np.int64, None),
                                     (np.float64, np.uint64)],
                         ids=lambda t: "none" if t is None else np.dtype(t).name)
@pytest.mark.parametrize("upper", [False, True], ids=["lower_cut", "upper_cut"])
def test_types_and_cuts(handle, kdt, vdt, upper):
    for n, nspl in [(1, 0), (1000, 1), (18433, 7), (300_001, 63), (1_000_003, 255)]:
        run(handle, kdt, vdt, n, nspl, upper, descending=False, seed=n + nspl)


@pytest.mark.parametrize("kdt", [np.uint32, np.int64, np.float32], ids=lambda d: np.dtype(d).name)
def test_descending(handle, kdt):
    for upper in (False, True):
        run(handle, kdt, np.uint32, 500_009, 31, upper, descending=True, seed=5)


@pytest.mark.parametrize("distinct", [1, 5, 300], ids=["equal", "five", "few"])
def test_skewed(handle, distinct):
    run(handle, np.uint32, np.uint32, 2_000_000, 64, False, False, seed=9, distinct=distinct)
    run(handle, np.uint32, np.uint32, 2_000_000, 64, True, False, seed=10, distinct=distinct)


def test_offsets_equal_bounds_of_sorted(handle):
    # On sorted input the multisplit offsets are the lower/upper-bound cuts.
    n = 1 << 20
    keys = np.sort(keys_of(np.uint32, n, 12, distinct=4000))
    spl = splitters_from(keys, 40, 13, False)
    for upper in (False, True):
        tk = dev(keys)
        ok = torch.empty_like(tk)
        offs = handle.multisplit(tk, spl, ok, upper=upper)
        side = "left" if upper else "right"
        # lower_bound cut of splitter s == number of keys < s == searchsorted left, and so on
        cuts = np.searchsorted(keys, spl, side="right" if upper else "left")
        assert np.array_equal(offs[1:-1].astype(np.int64), cuts), side
        assert_bits_equal(host(ok), keys, "sorted input stays in place")
        b = handle.bucket_bounds(tk, np.array([0, n], dtype=np.uint64), spl, upper=upper)
        assert np.array_equal(np.asarray(b).reshape(-1).astype(np.int64), offs.astype(np.int64))


def edge_values(kdt):
    d = np.dtype(kdt)
    if d.kind == "f":
        return np.array([-np.finfo(d).max, np.finfo(d).max], dtype=d)
    info = np.iinfo(d)
    return np.array([info.min, info.max], dtype=d)


@pytest.mark.parametrize("kdt", [np.uint32, np.int32, np.uint64, np.int64, np.float32, np.float64],
                         ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("upper", [False, True], ids=["lower_cut", "upper_cut"])
def test_bucket_table_edges(handle, kdt, upper):
    """Splitter sets that stress the bucket table: packed into a narrow range
    of a wide key spread (most table entries hold many splitters), at the
    type's extreme values, all equal, and far from most keys."""
    n = 400_003
    keys = keys_of(kdt, n, 31)
    lo, hi = edge_values(kdt)
    mid = np.sort(keys)[n // 2]
    near = np.sort(keys)[n // 2 - 40: n // 2 + 40]  # 80 consecutive distinct-ish keys
    cases = {
        "narrow": np.sort(near[::2]),
        "extremes": np.sort(np.array([lo, lo, mid, hi, hi], dtype=keys.dtype)),
        "all_equal": np.full(17, mid, dtype=keys.dtype),
        "far_low": np.sort(np.concatenate([np.full(3, lo, dtype=keys.dtype), near[:5]])),
        "full_255": np.sort(keys[(datagen.draws(77, 255) % np.uint64(n)).astype(np.int64)]),
    }
    # keys equal to the extremes and to the splitters, on purpose
    keys = keys.copy()
    keys[:4] = [lo, hi, mid, near[3]]
    for name, spl in cases.items():
        say(f"bucket table edge {name} kdt={np.dtype(kdt).name} upper={upper} nspl={spl.size}")
        check(handle, keys, np.uint32, spl, upper, descending=False)
        check(handle, keys, None, spl[::-1].copy(), upper, descending=True)


@pytest.mark.parametrize("kdt", [np.uint32, np.int64], ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("span", [2000, 6000, 12000, 30000])
@pytest.mark.parametrize("upper", [False, True], ids=["lower_cut", "upper_cut"])
def test_small_spans(handle, kdt, span, upper):
    """Dense keys around splitters whose span maps one to a few key images to
    a table entry (the counting read's refined entries then cover one or two
    images): every key value of the span and its surroundings occurs."""
    n = 200_001
    base = 1_000_000
    keys = (base - 300 + (datagen.draws(span, n) % np.uint64(span + 600)).astype(np.int64)).astype(kdt)
    spl = np.sort(base + (datagen.draws(span + 1, 40) % np.uint64(span + 1)).astype(np.int64)).astype(kdt)
    say(f"multisplit small span {span} kdt={np.dtype(kdt).name} upper={upper}")
    check(handle, keys, np.uint32, spl, upper, descending=False)
    check(handle, keys, None, spl[::-1].copy(), upper, descending=True)
