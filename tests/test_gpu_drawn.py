"""Seeded drawn configurations beyond the hand-picked ones: single-array
sorts (key type, payload type, direction, size, distribution), batches of
segments, multisplits, 2-way merges, gathers, host-pointer sorts and segment
plans with graph replays, every one against the oracle bit for bit."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import gpusort_amd as g
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


MERGE_SIZES = [0, 1, 5, 1000, 2047, 2048, 2049, 70001, 1 << 20]


def _merge_cases():
    return [(KEY_TYPES[x[0] % 6], bool(x[1] & 1), MERGE_SIZES[x[2] % 9], MERGE_SIZES[x[3] % 9], bool(x[4] & 1),
             x[5] % 9999) for x in _draws(7171, 8)]


@pytest.mark.parametrize("case", _merge_cases(), ids=lambda c: f"{_name(c[0])}-{'desc' if c[1] else 'asc'}-"
                         f"{c[2]}x{c[3]}{'-ties' if c[4] else ''}")
def test_drawn_merges(handle, case):
    kdt, desc, na, nb, ties, seed = case
    a = make_keys(kdt, na, seed=seed)
    b = make_keys(kdt, nb, seed=seed + 1)
    if ties and na:
        a = a[datagen.draws(seed + 2, na) % np.uint64(max(1, na // 20))]
        if nb:
            b = a[datagen.draws(seed + 3, nb) % np.uint64(na)]  # b's keys all occur in a
    a = oracle.std_sort(np.ascontiguousarray(a), desc)
    b = oracle.std_sort(np.ascontiguousarray(b), desc)
    say(f"drawn merge {case}")
    ok, ov = dev(np.zeros(na + nb, dtype=kdt)), dev(np.zeros(na + nb, dtype=np.uint32))
    handle.merge(dev(a), dev(b), ok, dev(np.arange(na, dtype=np.uint32)), dev(np.arange(na, na + nb, dtype=np.uint32)),
                 ov, descending=desc)
    assert_bits_equal(host(ok), oracle.merge(a, b, desc), "keys")
    cat = np.concatenate([a, b])
    order = np.argsort(-cat if np.dtype(kdt).kind == "f" and desc else (~cat if desc else cat), kind="stable")
    assert_bits_equal(host(ov), order.astype(np.uint32), "ties from the first run")


@pytest.mark.parametrize("case", [(np.uint16, 70001, 1000), (np.uint32, 1 << 20, 1 << 20), (np.uint64, 1000, 5),
                                  (np.uint32, 0, 7), (np.uint64, 1, 1), (np.uint16, 2049, 1 << 20)],
                         ids=lambda c: f"{_name(c[0])}-n{c[1]}-src{c[2]}")
def test_drawn_gathers(handle, case):
    et, n, nsrc = case
    src = (datagen.draws(nsrc, nsrc) >> np.uint64(8)).astype(et)
    idx = (datagen.draws(n + 3, n) % np.uint64(nsrc)).astype(np.int64)  # with repeats
    say(f"drawn gather {case}")
    dst = dev(np.zeros(n, dtype=et))
    handle.gather(dev(src), dev(idx), dst)
    assert np.array_equal(host(dst), src[idx])


@pytest.mark.parametrize("case", [(KEY_TYPES[x[0] % 6], bool(x[1] & 1), MERGE_SIZES[x[2] % 9], x[3] % 999)
                                  for x in _draws(8181, 6)],
                         ids=lambda c: f"{_name(c[0])}-{'desc' if c[1] else 'asc'}-n{c[2]}")
def test_drawn_host_pointer_sorts(handle, case):
    kdt, desc, n, seed = case
    keys = np.ascontiguousarray(make_keys(kdt, n, seed=seed))
    say(f"drawn host sort {case}")
    k = keys.copy()
    handle.sort_keys_host(k, descending=desc)
    assert_bits_equal(k, oracle.merge_sort(keys, descending=desc), "keys")
    vals = np.arange(n, dtype=np.uint32)
    k, v = keys.copy(), vals.copy()
    handle.sort_pairs_host(k, v, descending=desc)
    wk, wv = oracle.stable_sort_pairs(keys, vals, descending=desc)
    assert_bits_equal(k, wk, "pair keys")
    assert np.array_equal(v, wv)


@pytest.mark.parametrize("case", [(np.uint32, True, False, 500, 3000), (np.uint64, False, True, 10, 100000),
                                  (np.float32, True, True, 1, 0), (np.uint32, False, False, 500, 3)],
                         ids=lambda c: f"{_name(c[0])}-{'kv' if c[1] else 'keys'}-{'desc' if c[2] else 'asc'}"
                         f"-s{c[3]}-max{c[4]}")
def test_drawn_segment_plans(handle, case):
    """A plan run once, then replayed as a graph twice with eager sorts in
    between; segments include empty ones (all empty in one case)."""
    kdt, kv, desc, nseg, hi = case
    lengths = (datagen.draws(nseg + hi, nseg) % np.uint64(hi + 1)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint64)
    n = int(off[-1])
    keys = np.ascontiguousarray(make_keys(kdt, n, seed=nseg))
    vals = np.arange(n, dtype=np.uint32)
    if kv:
        wk, wv = oracle.seg_stable_sort_pairs(keys, vals, off.astype(np.int64), descending=desc)
    else:
        wk, wv = oracle.seg_sort(keys, off.astype(np.int64), descending=desc), None
    code = {np.dtype(np.uint32): g.U32, np.dtype(np.uint64): g.U64, np.dtype(np.float32): g.F32}[np.dtype(kdt)]
    plan = handle.seg_plan(off, code, g.U32 if kv else g.NONE, descending=desc)
    say(f"drawn plan {case} n={n}")
    for rep in range(3):
        tk, tv = dev(keys), dev(vals) if kv else None
        if rep == 0:
            plan.run(tk, tv)
        else:
            plan.launch_graph(tk, tv)
        torch.cuda.synchronize()
        assert_bits_equal(host(tk), wk, f"keys, replay {rep}")
        if kv:
            assert np.array_equal(host(tv), wv), f"payloads, replay {rep}"
        handle.sort_keys(dev(make_keys(np.uint32, 5000, seed=rep)))  # eager work between replays
    plan.close()
