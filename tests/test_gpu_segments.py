"""Batched segmented sort (gs_sort_segments) against the oracle's per-segment
sort, bit for bit: every size class (LDS tile, single radix tile, multi-tile
segmented onesweep) and its boundaries, all key types, payload widths and both
directions, skewed and constant keys, and a lognormal batch shaped like the
batched-arrays workload."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from gpusort_amd import datagen
from gpu_util import assert_bits_equal, dev, host, say

pytestmark = pytest.mark.gpu

# Around every class boundary: LDS tile caps 2048/4096/8192, onesweep tiles
# 3072/4096 (small), 9216 (64-bit keys with 8-byte payloads), 13312 (12-byte
# key + payload), 16384 (64-bit keys only), 19456 (32-bit keys with 4-byte
# payloads) and 28672 (32-bit keys only).
BOUNDARY_LENGTHS = [0, 1, 2, 3, 2047, 2048, 2049, 4095, 4096, 4097, 6143, 6144, 6145, 8191, 8192, 8193,
                    9215, 9216, 9217, 10239, 10240, 10241, 12288, 12289, 13311, 13312, 13313, 16383, 16384, 16385, 18431, 18432, 18433, 19455, 19456, 19457, 20480, 20481,
                    24575, 24576, 24577, 26624, 26625, 28671, 28672, 28673, 30720, 32768, 36864, 36865, 1, 0, 49153, 50000,
                    100000, 262147, 5]


def keys_of(dtype, n, seed, distinct=None):
    d = np.dtype(dtype)
    r = datagen.draws(seed, n)
    if distinct:
        r = datagen.draws(seed + 1, distinct)[r % np.uint64(distinct)]
    if d.kind == "f":
        x = (r >> np.uint64(40)).astype(np.float64) - 2.0**23
        return x.astype(d)
    if d.itemsize == 4:
        return (r >> np.uint64(32)).astype(np.uint32).view(d)
    return r.view(d)


def offsets_of(lengths) -> np.ndarray:
    return np.concatenate([[0], np.cumsum(np.asarray(lengths, dtype=np.uint64))]).astype(np.uint64)


def run_case(handle, kdt, vdt, lengths, seed, descending, distinct=None):
    off = offsets_of(lengths)
    n = int(off[-1])
    keys = keys_of(kdt, n, seed, distinct)
    vals = None if vdt is None else np.arange(n, dtype=np.uint64).astype(vdt)
    say(f"segments kdt={np.dtype(kdt).name} vdt={None if vdt is None else np.dtype(vdt).name} "
        f"nseg={len(lengths)} n={n} desc={descending} seed={seed}")
    tk = dev(keys)
    tv = None if vals is None else dev(vals)
    handle.sort_segments(tk, dev(off), tv, descending=descending)
    off64 = off.astype(np.int64)
    if vals is None:
        want = oracle.seg_sort(keys, off64, descending)
        assert_bits_equal(host(tk), want, "keys")
    else:
        wk, wv = oracle.seg_stable_sort_pairs(keys, vals, off64, descending)
        assert_bits_equal(host(tk), wk, "keys")
        assert_bits_equal(host(tv), wv, "payloads")


TYPES = [
    (np.uint32, None), (np.uint32, np.uint32), (np.int32, np.uint64), (np.float32, np.uint32),
    (np.uint64, None), (np.uint64, np.uint32), (np.int64, np.uint64), (np.float64, np.uint32),
]


@pytest.mark.parametrize("kdt,vdt", TYPES,
                         ids=lambda t: "none" if t is None else np.dtype(t).name)
def test_boundary_lengths(handle, kdt, vdt):
    run_case(handle, kdt, vdt, BOUNDARY_LENGTHS, 21, descending=False)


@pytest.mark.parametrize("kdt,vdt", [(np.uint32, np.uint32), (np.int64, None), (np.float32, np.uint64)],
                         ids=lambda t: "none" if t is None else np.dtype(t).name)
def test_boundary_lengths_descending(handle, kdt, vdt):
    run_case(handle, kdt, vdt, BOUNDARY_LENGTHS[::-1], 22, descending=True)


@pytest.mark.parametrize("distinct", [1, 3, 200], ids=["equal", "three", "few"])
def test_skewed_keys(handle, distinct):
    # Ties everywhere: stability across the tiles of a segment is what is checked.
    run_case(handle, np.uint32, np.uint32, BOUNDARY_LENGTHS, 23, descending=False, distinct=distinct)


def test_lognormal_batch(handle):
    lengths = datagen.lognormal_lengths(5, 3000, median=2048, sigma=2.0, lo=1, hi=1 << 20)
    run_case(handle, np.uint32, np.uint32, lengths, 24, descending=False)


def test_many_tiny_segments(handle):
    lengths = datagen.lognormal_lengths(6, 40000, median=40, sigma=1.0, lo=0, hi=5000)
    run_case(handle, np.uint64, np.uint32, lengths, 25, descending=True)


def test_one_large_segment_equals_single_sort(handle):
    # The segmented path over one segment must equal the single-array sort.
    n = 3_000_017
    keys = keys_of(np.uint32, n, 26)
    vals = np.arange(n, dtype=np.uint32)
    a_k, a_v = dev(keys), dev(vals)
    b_k, b_v = dev(keys), dev(vals)
    say(f"one segment n={n}")
    handle.sort_segments(a_k, dev(offsets_of([n])), a_v)
    handle.sort_pairs(b_k, b_v)
    assert_bits_equal(host(a_k), host(b_k), "keys")
    assert_bits_equal(host(a_v), host(b_v), "payloads")


def test_repeat_calls_shrinking(handle):
    # Workspace and maps are reused: a second, smaller call must not see the first.
    run_case(handle, np.uint32, np.uint32, [70000, 20000, 5000, 100], 27, descending=False)
    run_case(handle, np.uint32, np.uint32, [17000, 3], 28, descending=False)


U32, U64, NONE = 1, 4, 255


@pytest.mark.parametrize("small_max,medium_max,geometry",
                         [(4096, 0, 0), (1024, 1 << 16, 0), (1024, 1 << 16, 1), (1024, 1 << 16, 2),
                          (256, 1 << 20, 1), (256, 1 << 20, 2)],
                         ids=["default", "small-tiles-64K", "half-tiles-64K", "wide-tiles-64K", "half-tiles-1M",
                              "wide-tiles-1M"])
def test_size_class_configs_agree(handle, small_max, medium_max, geometry):
    # Every size-class configuration and small-tile geometry sorts identically.
    handle.set_seg_config(U32, U32, small_max, medium_max, geometry)
    try:
        assert handle.seg_config(U32, U32) == {"small_max": small_max, "medium_max": medium_max,
                                               "medium_geometry": geometry}
        run_case(handle, np.uint32, np.uint32, BOUNDARY_LENGTHS, 31, descending=False)
        lengths = datagen.lognormal_lengths(7, 2000, median=2048, sigma=2.0, lo=1, hi=1 << 20)
        run_case(handle, np.uint32, np.uint32, lengths, 32, descending=True)
    finally:
        handle.set_seg_config(U32, U32, 4096, 0)


# The small-tile geometries of the other widths (x 12 and x 6 keys per
# thread when a key and payload take more than 8 bytes; keys alone).
@pytest.mark.parametrize("kdt,vdt,kcode,vcode", [(np.uint64, np.uint64, U64, U64), (np.uint64, np.uint32, U64, U32),
                                                 (np.uint32, None, U32, NONE)],
                         ids=["u64+u64", "u64+u32", "u32"])
@pytest.mark.parametrize("geometry", [1, 2])
def test_small_tile_geometries_every_width(handle, kdt, vdt, kcode, vcode, geometry):
    before = handle.seg_config(kcode, vcode)
    handle.set_seg_config(kcode, vcode, 512, 1 << 16, geometry)
    try:
        lengths = datagen.lognormal_lengths(11, 1200, median=2048, sigma=2.0, lo=1, hi=1 << 18)
        run_case(handle, kdt, vdt, lengths, 35 + geometry, descending=geometry == 2)
    finally:
        handle.set_seg_config(kcode, vcode, before["small_max"], before["medium_max"], before["medium_geometry"])


def test_autotune_picks_the_fastest_candidate(handle):
    say("autotune u32+u32 on 2^22 keys")
    best, table = handle.autotune_segments(U32, U32, total_keys=1 << 22)
    try:
        # 3 LDS-tile limits x (no small-tile class, or 2 limits x 3 geometries)
        assert len(table) == 21 and all(r["ms"] > 0 for r in table)
        fastest = min(table, key=lambda r: r["ms"])
        assert best == {"small_max": fastest["small_max"], "medium_max": fastest["medium_max"],
                        "medium_geometry": fastest["medium_geometry"]}
        assert handle.seg_config(U32, U32) == best
        lengths = datagen.lognormal_lengths(8, 1500, median=2048, sigma=2.0, lo=1, hi=1 << 20)
        run_case(handle, np.uint32, np.uint32, lengths, 33, descending=False)
    finally:
        handle.set_seg_config(U32, U32, 4096, 0)


def test_plan_run_and_graph_replays(handle):
    import torch

    lengths = datagen.lognormal_lengths(9, 1500, median=2048, sigma=2.0, lo=1, hi=1 << 20)
    off = offsets_of(lengths)
    n = int(off[-1])
    keys = keys_of(np.uint32, n, 34)
    vals = np.arange(n, dtype=np.uint32)
    wk, wv = oracle.seg_stable_sort_pairs(keys, vals, off.astype(np.int64), False)
    plan = handle.seg_plan(off, U32, U32)
    say(f"plan run n={n}")
    tk, tv = dev(keys), dev(vals)
    plan.run(tk, tv)
    torch.cuda.synchronize()
    assert_bits_equal(host(tk), wk, "plan keys")
    assert_bits_equal(host(tv), wv, "plan payloads")
    src_k, src_v = dev(keys), dev(vals)
    gk, gv = torch.empty_like(src_k), torch.empty_like(src_v)
    other = dev(keys_of(np.uint32, 1 << 16, 35))
    for replay in range(3):
        gk.copy_(src_k)
        gv.copy_(src_v)
        torch.cuda.synchronize()
        say(f"graph replay {replay}")
        plan.launch_graph(gk, gv)
        torch.cuda.synchronize()
        assert_bits_equal(host(gk), wk, f"graph replay {replay} keys")
        assert_bits_equal(host(gv), wv, f"graph replay {replay} payloads")
        handle.sort_keys(other)  # eager work between replays
        torch.cuda.synchronize()
    # new buffers: the plan captures a new graph
    hk, hv = dev(keys), dev(vals)
    plan.launch_graph(hk, hv)
    torch.cuda.synchronize()
    assert_bits_equal(host(hk), wk, "recaptured keys")
    plan.close()
