"""Parity of the device sort (gs_sort_keys / gs_sort_pairs) with the oracle.

Keys-only results must equal the oracle's restatement of OmpUtils::MergeSort
bit for bit; key-value results must equal the stable sort by key that
thrust::sort_by_key gives (src/cuda_utils.cu:477), checked against both the
oracle and numpy's stable argsort.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from gpusort_amd import datagen
from tests.gpu_util import assert_bits_equal, dev, host, say, stable_order

pytestmark = pytest.mark.gpu

KEY_DTYPES = [np.uint32, np.int32, np.float32, np.uint64, np.int64, np.float64]

# Sizes around the block-sort cap (8192 / 4096 / 2048) and the onesweep tiles
# (at these sizes 16384 keys-only u32, 10240 u32 + 32-bit payload and 8192
# keys-only u64, the mid-size tiles; 13312 with 12-byte key + payload, 9216
# u64 + 64-bit payload; 18432, 10240 and 6144 for multisplit), plus
# multi-tile and ragged lengths.  The large single-array tiles are covered for
# every key type and direction by test_large_tiles_every_type below.
SIZES = [0, 1, 2, 3, 63, 64, 65, 1000, 2047, 2048, 2049, 4095, 4096, 4097, 6143, 6144, 6145, 8191, 8192, 8193,
         9215, 9216, 9217, 10239, 10240, 10241, 12287, 12288, 12289, 13311, 13312, 13313, 16383, 16384, 16385, 18431, 18432, 18433, 19455, 19456, 19457, 24575, 24576, 24577, 26625, 28671, 28672, 28673, 30719, 30720, 30721, 49153, 61441, 100_003, 1 << 20]


def make_keys(dtype, n: int, seed: int) -> np.ndarray:
    d = np.dtype(dtype)
    r = datagen.draws(seed, n)
    if d == np.float32:
        f = datagen.normal_f32(seed, n)
        return f
    if d == np.float64:
        z = datagen.std_normal(seed, n) * 1e6
        z[z == 0] = 0.0
        return z.astype(np.float64)
    if d.itemsize == 4:
        return (r >> np.uint64(32)).astype(np.uint32).view(d)
    return r.view(d)


@pytest.mark.parametrize("dtype", KEY_DTYPES, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("descending", [False, True], ids=["asc", "desc"])
def test_sort_keys_sizes(handle, dtype, descending):
    for n in SIZES:
        keys = make_keys(dtype, n, seed=n + 7)
        say(f"sort_keys dtype={np.dtype(dtype).name} n={n} desc={descending}")
        t = dev(keys)
        handle.sort_keys(t, descending=descending)
        got = host(t)
        want = oracle.merge_sort(keys, descending=descending)
        assert_bits_equal(got, want, f"n={n}")


@pytest.mark.parametrize("kdtype", KEY_DTYPES, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("vdtype", [np.uint32, np.uint64], ids=["v32", "v64"])
@pytest.mark.parametrize("descending", [False, True], ids=["asc", "desc"])
def test_sort_pairs_stable(handle, kdtype, vdtype, descending):
    for n in [0, 1, 5, 2048, 2049, 4097, 6145, 8193, 50_001, 1 << 19]:
        # Few distinct keys so that stability is exercised.
        keys = make_keys(kdtype, n, seed=3 * n + 1)
        if n > 10:
            keys = keys[datagen.draws(99, n) % np.uint64(max(1, n // 16))]
        vals = np.arange(n, dtype=vdtype)
        say(f"sort_pairs k={np.dtype(kdtype).name} v={np.dtype(vdtype).name} n={n} desc={descending}")
        tk, tv = dev(keys), dev(vals)
        handle.sort_pairs(tk, tv, descending=descending)
        gk, gv = host(tk), host(tv)
        wk, wv = oracle.stable_sort_pairs(keys, vals, descending=descending)
        assert_bits_equal(gk, wk, f"keys n={n}")
        assert_bits_equal(gv, wv, f"values n={n}")
        perm = stable_order(keys, descending)
        assert np.array_equal(gv, vals[perm])


# Single-array sorts take smaller tiles up to 2^23 keys (32-bit keys alone)
# and 2^21 keys (32-bit keys with 32-bit payloads); these sizes sit on both
# sides of each switch, with partial last tiles in the larger geometry.
@pytest.mark.parametrize("n", [(1 << 21), (1 << 21) + 1, (1 << 21) + 19457, (1 << 23), (1 << 23) + 1,
                               (1 << 23) + 30721])
def test_tile_geometry_thresholds(handle, n):
    k64 = make_keys(np.uint64, min(n, (1 << 21) + 16385), seed=n + 13)  # 64-bit keys switch at 2^21
    say(f"sort_keys u64 n={k64.size}")
    t64 = dev(k64)
    handle.sort_keys(t64)
    assert_bits_equal(host(t64), oracle.merge_sort(k64), f"u64 keys n={k64.size}")
    keys = make_keys(np.uint32, n, seed=n + 11)
    say(f"sort_keys u32 n={n}")
    t = dev(keys)
    handle.sort_keys(t)
    assert_bits_equal(host(t), oracle.merge_sort(keys), f"keys n={n}")
    kk = keys[datagen.draws(5, n) % np.uint64(n // 16)]
    vals = np.arange(n, dtype=np.uint32)
    say(f"sort_pairs u32+u32 n={n}")
    tk, tv = dev(kk), dev(vals)
    handle.sort_pairs(tk, tv)
    perm = stable_order(kk, False)
    assert_bits_equal(host(tk), kk[perm], f"pair keys n={n}")
    assert_bits_equal(host(tv), vals[perm], f"pair values n={n}")


# The large single-array tiles for every key type and direction (the
# transformed-key kernels that signed, float and descending sorts use), one
# size just above each switch: keys alone (32-bit keys past 2^23, 64-bit past
# 2^21) and 32-bit keys with 32-bit payloads past 2^21.
@pytest.mark.parametrize("dtype", KEY_DTYPES, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("descending", [False, True], ids=["asc", "desc"])
def test_large_tiles_every_type(handle, dtype, descending):
    d = np.dtype(dtype)
    n = (1 << 23) + 30721 if d.itemsize == 4 else (1 << 21) + 16385
    keys = make_keys(dtype, n, seed=n + 3 * d.itemsize + int(descending))
    say(f"large tile keys {d.name} n={n} desc={descending}")
    t = dev(keys)
    handle.sort_keys(t, descending=descending)
    assert_bits_equal(host(t), oracle.merge_sort(keys, descending=descending), f"keys {d.name} n={n}")
    nk = (1 << 21) + 19457
    kk = make_keys(dtype, nk, seed=nk + d.itemsize)
    kk = kk[datagen.draws(17, nk) % np.uint64(nk // 16)]
    vals = np.arange(nk, dtype=np.uint32)
    say(f"large tile pairs {d.name}+u32 n={nk} desc={descending}")
    tk, tv = dev(kk), dev(vals)
    handle.sort_pairs(tk, tv, descending=descending)
    wk, wv = oracle.stable_sort_pairs(kk, vals, descending=descending)
    assert_bits_equal(host(tk), wk, f"pair keys {d.name} n={nk}")
    assert_bits_equal(host(tv), wv, f"pair values {d.name} n={nk}")


# Skewed keys alone on the large tile: whole waves of one digit take the
# uniform-wave ranking branch.
@pytest.mark.parametrize("dist", ["zipf", "equal", "sorted"])
@pytest.mark.parametrize("descending", [False, True], ids=["asc", "desc"])
def test_large_tile_skewed_keys(handle, dist, descending):
    n = (1 << 23) + 30721
    keys = datagen.keys(dist, 21, n)
    say(f"large tile skewed keys {dist} n={n} desc={descending}")
    t = dev(keys)
    handle.sort_keys(t, descending=descending)
    want = np.sort(keys)
    assert_bits_equal(host(t), want[::-1].copy() if descending else want, f"{dist} keys n={n}")


@pytest.mark.parametrize("dist", ["uniform", "zipf", "sorted", "equal"])
def test_distributions_kv(handle, dist):
    n = 3_000_017
    keys = datagen.keys(dist, 11, n)
    vals = np.arange(n, dtype=np.uint32)
    say(f"distribution {dist} n={n}")
    tk, tv = dev(keys), dev(vals)
    handle.sort_pairs(tk, tv)
    gk, gv = host(tk), host(tv)
    perm = np.argsort(keys, kind="stable")
    assert_bits_equal(gk, keys[perm], "keys")
    assert_bits_equal(gv, perm.astype(np.uint32), "payload permutation")


def test_float_special_order(handle):
    # -0.0 orders before +0.0 under the order-preserving bit flip; the
    # reference's std::less<float> leaves their relative order unspecified.
    keys = np.array([3.5, -0.0, 0.0, -1e-30, 1e-30, -3.5, np.finfo(np.float32).max, -np.finfo(np.float32).max],
                    dtype=np.float32)
    t = dev(keys)
    handle.sort_keys(t)
    got = host(t)
    assert got.tolist() == sorted(keys.tolist())
    assert np.signbit(got[3]) and not np.signbit(got[4])


def test_host_entries(handle):
    n = 70_001
    keys = datagen.uniform_u32(5, n)
    want = oracle.merge_sort(keys)
    k = keys.copy()
    handle.sort_keys_host(k)
    assert_bits_equal(k, want, "host keys")
    keys = datagen.uniform_u64(6, n).view(np.int64)
    vals = np.arange(n, dtype=np.uint32)
    pairs = np.zeros(n, dtype=np.dtype([("k", np.int64), ("v", np.uint32), ("pad", np.uint32)]))
    k, v = keys.copy(), vals.copy()
    handle.sort_pairs_host(k, v, pairs_out=pairs, pair_stride=16, val_offset=8)
    wk, wv = oracle.stable_sort_pairs(keys, vals)
    assert_bits_equal(k, wk, "host pair keys")
    assert_bits_equal(v, wv, "host pair values")
    assert np.array_equal(pairs["k"], wk) and np.array_equal(pairs["v"], wv)


def test_repeat_calls_reuse_arena(handle):
    keys = datagen.uniform_u32(8, 1 << 20)
    t = dev(keys)
    handle.sort_keys(t)
    before = handle.arena_bytes()
    for s in range(3):
        k2 = datagen.uniform_u32(100 + s, 1 << 20)
        t = dev(k2)
        handle.sort_keys(t)
        assert_bits_equal(host(t), np.sort(k2), f"repeat {s}")
    assert handle.arena_bytes() == before


def test_tile_order_choices(handle):
    """gs_set_tile_order: arrival-order tiles on the null stream and grouped
    tiles on a side stream give the same results as the default choice;
    an unknown order is refused."""
    import torch

    import gpusort_amd as g

    n = (1 << 22) + 19457
    keys = make_keys(np.uint32, n, seed=77)
    vals = np.arange(n, dtype=np.uint32)
    wk, wv = oracle.stable_sort_pairs(keys, vals)
    side = torch.cuda.Stream()
    try:
        say(f"tile order arrival, null stream, pairs n={n}")
        handle.set_tile_order(g.TILES_ARRIVAL)
        tk, tv = dev(keys), dev(vals)
        handle.sort_pairs(tk, tv)
        assert_bits_equal(host(tk), wk, "arrival keys")
        assert_bits_equal(host(tv), wv, "arrival values")
        say("tile order arrival, segments and multisplit")
        off = np.array([0, 5000, 5000, n // 2, n], dtype=np.uint64)
        tk, tv = dev(keys), dev(vals)
        handle.sort_segments(tk, torch.from_numpy(off.view(np.int64)).cuda(), tv)
        gk, gv = host(tk), host(tv)
        for a, b in zip(off[:-1].astype(np.int64), off[1:].astype(np.int64)):
            sk, sv = oracle.stable_sort_pairs(keys[a:b], vals[a:b])
            assert_bits_equal(gk[a:b], sk, f"segment keys [{a},{b})")
            assert_bits_equal(gv[a:b], sv, f"segment values [{a},{b})")
        spl = np.sort(keys[:: n // 33])[1:33]
        ok, ov = torch.empty_like(dev(keys)), torch.empty_like(dev(vals))
        offs = handle.multisplit(dev(keys), spl, ok, dev(vals), ov)
        idx, want_offs = oracle.multisplit(keys, spl, False, False)
        assert np.array_equal(offs.astype(np.int64), want_offs)
        assert_bits_equal(host(ok), keys[idx], "multisplit keys")
        say(f"tile order grouped, side stream, keys n={n}")
        handle.set_tile_order(g.TILES_GROUPED)
        handle.set_stream(side.cuda_stream)
        with torch.cuda.stream(side):
            t = dev(keys)
            handle.sort_keys(t)
            side.synchronize()
        assert_bits_equal(host(t), oracle.merge_sort(keys), "grouped keys on a side stream")
        with pytest.raises(Exception):
            handle.set_tile_order(7)
    finally:
        handle.set_stream(0)
        handle.set_tile_order(g.TILES_AUTO)
