"""Seeded drawn configurations beyond the hand-picked ones: single-array
sorts (key type, payload type, direction, size, distribution), batches of
segments and multisplits, every one against the oracle bit for bit."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from gpusort_amd import datagen
from tests.gpu_util import assert_bits_equal, dev, host, say
from tests.test_gpu_multisplit import run as multisplit_run
from tests.test_gpu_sort import make_keys

pytestmark = pytest.mark.gpu

KEY_TYPES = [np.uint32, np.int32, np.float32, np.uint64, np.int64, np.float64]
VAL_TYPES = [None, np.uint32, np.uint64]
SIZES = [1, 7, 1000, 4097, 19456, 19457, 65536, 1 << 20, (1 << 21) + 5, (1 << 22) + 999]


def _draws(seed, count, width=8):
    d = datagen.draws(seed, width * count)
    return [[int(v) for v in d[width * c: width * c + width]] for c in range(count)]


def _sort_cases():
    out = []
    for x in _draws(4242, 24):
        out.append((KEY_TYPES[x[0] % 6], VAL_TYPES[x[1] % 3], bool(x[2] & 1), SIZES[x[3] % len(SIZES)], x[4] % 4,
                    x[5] % 100000))
    return out


def _name(t):
    return "none" if t is None else np.dtype(t).name


@pytest.mark.parametrize("case", _sort_cases(), ids=lambda c: f"{_name(c[0])}-{_name(c[1])}-"
                         f"{'desc' if c[2] else 'asc'}-n{c[3]}-dist{c[4]}")
def test_drawn_sorts(handle, case):
    kdt, vdt, desc, n, dist, seed = case
    keys = make_keys(kdt, n, seed=seed)
    if dist == 1:  # about 50 keys per distinct value
        keys = keys[datagen.draws(seed + 1, n) % np.uint64(max(1, n // 50))]
    elif dist == 2:  # already sorted
        keys = np.sort(keys)
    elif dist == 3:  # all equal
        keys = np.full(n, keys[0], dtype=keys.dtype)
    keys = np.ascontiguousarray(keys)
    say(f"drawn sort {case}")
    tk = dev(keys)
    if vdt is None:
        handle.sort_keys(tk, descending=desc)
        assert_bits_equal(host(tk), oracle.merge_sort(keys, descending=desc), "keys")
    else:
        vals = np.arange(n, dtype=np.uint64).astype(vdt)
        tv = dev(vals)
        handle.sort_pairs(tk, tv, descending=desc)
        wk, wv = oracle.stable_sort_pairs(keys, vals, descending=desc)
        assert_bits_equal(host(tk), wk, "keys")
        assert_bits_equal(host(tv), wv, "payloads")


def _segment_cases():
    out = []
    for x in _draws(5151, 12):
        out.append((KEY_TYPES[x[0] % 6], bool(x[1] & 1), bool(x[2] & 1), [1, 3, 100, 2000][x[3] % 4],
                    [1, 50, 5000, 70000, 1 << 18][x[4] % 5], x[5] % 1000, bool(x[6] & 1)))
    return out


@pytest.mark.parametrize("case", _segment_cases(), ids=lambda c: f"{_name(c[0])}-{'kv' if c[1] else 'keys'}-"
                         f"{'desc' if c[2] else 'asc'}-s{c[3]}-max{c[4]}{'-ties' if c[6] else ''}")
def test_drawn_segment_batches(handle, case):
    kdt, kv, desc, nseg, hi, seed, ties = case
    lengths = (datagen.draws(seed, nseg) % np.uint64(hi + 1)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint64)
    n = int(off[-1])
    keys = make_keys(kdt, n, seed=seed + 1)
    if ties and n:
        keys = keys[datagen.draws(seed + 2, n) % np.uint64(max(1, n // 100))]
    keys = np.ascontiguousarray(keys)
    say(f"drawn segments {case} n={n}")
    tk = dev(keys)
    toff = torch.from_numpy(off.view(np.int64)).cuda()
    if kv:
        vals = np.arange(n, dtype=np.uint32)
        tv = dev(vals)
        handle.sort_segments(tk, toff, tv, descending=desc)
        wk, wv = oracle.seg_stable_sort_pairs(keys, vals, off.astype(np.int64), descending=desc)
        assert_bits_equal(host(tk), wk, "keys")
        assert_bits_equal(host(tv), wv, "payloads")
    else:
        handle.sort_segments(tk, toff, None, descending=desc)
        assert_bits_equal(host(tk), oracle.seg_sort(keys, off.astype(np.int64), descending=desc), "keys")


def _multisplit_cases():
    out = []
    for x in _draws(6161, 12):
        out.append((KEY_TYPES[x[0] % 6], VAL_TYPES[x[1] % 3], bool(x[2] & 1), bool(x[3] & 1),
                    [1, 1000, 77777, 1 << 20, 3000001][x[4] % 5], [0, 1, 7, 63, 200, 255][x[5] % 6], x[6] % 100000,
                    None if x[7] & 1 else 37))
    return out


@pytest.mark.parametrize("case", _multisplit_cases(), ids=lambda c: f"{_name(c[0])}-{_name(c[1])}-"
                         f"{'desc' if c[2] else 'asc'}-{'upper' if c[3] else 'lower'}-n{c[4]}-s{c[5]}"
                         f"{'-d37' if c[7] else ''}")
def test_drawn_multisplits(handle, case):
    kdt, vdt, desc, upper, n, nspl, seed, distinct = case
    multisplit_run(handle, kdt, vdt, n, nspl, upper, desc, seed=seed, distinct=distinct)
