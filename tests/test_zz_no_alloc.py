"""gs_sort_pairs_noalloc (the sort with the library's fixed device workspace)
against gs_sort_pairs / gs_sort_keys on the same inputs, bit for bit, through
the C ABI: two calls in a row per case, the second smaller, on a fresh handle
whose arena must stay empty (nothing allocated)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import gpusort_amd as g
from gpusort_amd import datagen
from tests.gpu_util import assert_bits_equal, dev, host, say

pytestmark = pytest.mark.gpu

# (key dtype, payload dtype or None, descending, first n, second n): the
# second call is smaller and follows the first on the same array.
CASES = [
    (np.uint32, np.uint32, False, (1 << 23) - 5, 1_000_003),  # largest u32+u32 that fits, then mid-size
    (np.uint32, None, False, 1 << 24, 3_000_017),               # keys alone, large tiles then mid-size
    (np.float32, np.uint32, True, 2_500_001, 40_003),
    (np.int64, np.uint64, False, 1 << 22, 9_999),
    (np.uint64, None, True, 5_000_011, 8_193),
    (np.int32, np.uint64, True, 100_000, 4_000),               # second call in one workgroup (no workspace)
]


def _keys(dtype, n, seed):
    d = np.dtype(dtype)
    r = datagen.draws(seed, n)
    if d == np.float32:
        return datagen.normal_f32(seed, n)
    if d.itemsize == 4:
        # few distinct keys, so payload order (stability) is checked
        return ((r >> np.uint64(40)).astype(np.uint32)).view(d)
    return (r >> np.uint64(20)).view(d)


def _call(lib, h, kdt, tk, vdt, tv, n, desc):
    return lib.gs_sort_pairs_noalloc(h, kdt, tk.data_ptr(), vdt, None if tv is None else tv.data_ptr(), n, int(desc))


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{np.dtype(c[0]).name}-{'none' if c[1] is None else np.dtype(c[1]).name}-{'desc' if c[2] else 'asc'}")
def test_noalloc_matches_sort_pairs(case):
    kdtype, vdtype, desc, n1, n2 = case
    lib = g.load_library()
    fresh = g.Handle(0)
    ref = g.Handle(0)
    try:
        kdt = g.np_dtype_code(np.zeros(1, kdtype))
        vdt = g.NONE if vdtype is None else g.np_dtype_code(np.zeros(1, vdtype))
        big_k = _keys(kdtype, n1, seed=n1 + 11)
        big_v = None if vdtype is None else np.arange(n1, dtype=vdtype)[::-1].copy()
        tk = dev(big_k)
        tv = None if big_v is None else dev(big_v)
        for call, n in enumerate((n1, n2)):
            say(f"noalloc call {call}: k={np.dtype(kdtype).name} v={vdtype} n={n} desc={desc}")
            a_k = tk[:n].clone()
            a_v = None if tv is None else tv[:n].clone()
            st = _call(lib, fresh.h, kdt, a_k, vdt, a_v, n, desc)
            assert st == 0, lib.gs_last_error()
            b_k = tk[:n].clone()
            b_v = None if tv is None else tv[:n].clone()
            if b_v is None:
                st = lib.gs_sort_keys(ref.h, kdt, b_k.data_ptr(), n, int(desc))
            else:
                st = lib.gs_sort_pairs(ref.h, kdt, b_k.data_ptr(), vdt, b_v.data_ptr(), n, int(desc))
            assert st == 0, lib.gs_last_error()
            assert lib.gs_synchronize(fresh.h) == 0 and lib.gs_synchronize(ref.h) == 0
            assert_bits_equal(host(a_k), host(b_k), f"keys call {call} n={n}")
            if a_v is not None:
                assert_bits_equal(host(a_v), host(b_v), f"values call {call} n={n}")
            # the keys really are sorted (the reference entry is not trusted alone)
            hk = host(a_k)
            if n > 1:
                assert (hk[1:] <= hk[:-1]).all() if desc else (hk[1:] >= hk[:-1]).all()
        assert fresh.arena_bytes() == 0, "the no-allocation entry grew the handle's arena"
    finally:
        fresh.close()
        ref.close()


def test_noalloc_rejects_too_large():
    lib = g.load_library()
    h = g.Handle(0)
    try:
        n = (1 << 23) + (1 << 20)  # u32 + u32: 72 MiB of ping-pong alone
        tk = dev(np.zeros(8, np.uint32))  # never touched: the size check comes first
        say(f"noalloc reject n={n}")
        st = lib.gs_sort_pairs_noalloc(h.h, g.U32, tk.data_ptr(), g.U32, tk.data_ptr(), n, 0)
        assert st == 1
        assert b"fixed scratch array" in lib.gs_last_error()
        assert h.arena_bytes() == 0
        # null pointers and bad types are refused before any device work
        assert lib.gs_sort_pairs_noalloc(h.h, 77, tk.data_ptr(), g.NONE, None, 10, 0) == 1
        assert lib.gs_sort_pairs_noalloc(h.h, g.U32, None, g.NONE, None, 10, 0) == 1
        assert lib.gs_sort_pairs_noalloc(None, g.U32, tk.data_ptr(), g.NONE, None, 10, 0) == 1
    finally:
        h.close()
