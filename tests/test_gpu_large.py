"""Sizes past the 32-bit paths: a sort of 2^30 + k keys (64-bit look-back
words and positions) and a segmented batch of more than 2^32 keys in total
(64-bit positions in the segmented passes).  Checked by size-independent
properties: order, order-independent checksums of the multiset, the payload
mapping input to output, and stability."""
from __future__ import annotations

import pytest
import torch

import gpusort_amd as g
from gpu_util import say

pytestmark = pytest.mark.gpu

MASK32 = 0xFFFFFFFF


def u32_as_i64(t):
    return t.view(torch.int32).to(torch.int64) & MASK32


def checksums(k64):
    # sum and sum of squares (wrapping int64) of the key values
    return int(k64.sum()), int((k64 * k64).sum())


def test_sort_pairs_past_2p30(handle):
    n = (1 << 30) + 4099
    keys = torch.empty(n, dtype=torch.uint32, device="cuda")
    handle.generate(keys, g.GEN_UNIFORM_U32, 41)
    vals = torch.empty(n, dtype=torch.uint32, device="cuda")
    handle.generate(vals, g.GEN_IOTA_U32, 0)
    k0 = u32_as_i64(keys)
    want = checksums(k0)
    say(f"sort_pairs n={n}")
    handle.sort_pairs(keys, vals)
    torch.cuda.synchronize()
    k = u32_as_i64(keys)
    assert bool((k[1:] >= k[:-1]).all())
    assert checksums(k) == want
    idx = u32_as_i64(vals)
    assert bool((k0[idx] == k).all())
    assert bool(((idx[1:] > idx[:-1]) | (k[1:] != k[:-1])).all())
    assert int(idx.sum()) == n * (n - 1) // 2


def test_segments_past_2p32_total(handle):
    # five segments just under 2^30 keys each (segmented radix), > 2^32 in all
    seg = (1 << 30) - 5
    lengths = [seg, seg, seg, seg, 3 * (1 << 20) + 7]
    off = [0]
    for ln in lengths:
        off.append(off[-1] + ln)
    n = off[-1]
    assert n > (1 << 32)
    keys = torch.empty(n, dtype=torch.uint32, device="cuda")
    handle.generate(keys, g.GEN_UNIFORM_U32, 42)
    want = [checksums(u32_as_i64(keys[off[s]:off[s + 1]])) for s in range(len(lengths))]
    d_off = torch.tensor(off, dtype=torch.int64, device="cuda")
    say(f"sort_segments n={n} segments={len(lengths)}")
    handle.sort_segments(keys, d_off)
    torch.cuda.synchronize()
    for s in range(len(lengths)):
        k = u32_as_i64(keys[off[s]:off[s + 1]])
        assert bool((k[1:] >= k[:-1]).all()), f"segment {s} not sorted"
        assert checksums(k) == want[s], f"segment {s} lost keys"
