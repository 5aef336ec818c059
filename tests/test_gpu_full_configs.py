"""The benchmark configurations at their full sizes, as GPU tests (the same
inputs bench.py measures):

- a single-rank 2^24 u32 sort against the oracle's OmpUtils::MergeSort;
- a stable 2^28 u32 + u32 sort, by size-independent properties (order, the
  payload maps input to output, ties in input order);
- 64 virtual ranks x 2^21 u64 keys + u32 payloads, k = 8 HykSort with 32
  samples per rank, against the oracle's re-run of the same rounds (keys,
  payloads, splitters, bucket sizes, final rank sizes); the same shape at
  2^19 keys per rank with 1000 distinct keys and random payloads;
- 100,000 arrays of lognormal lengths (median 2048, sigma 2, at most 2^22) in
  one batched call, by per-array properties;
- served requests of log-uniform sizes up to 2^28 and mixed types, each
  against the same input sorted alone.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import gpusort_amd as g
import oracle
from gpu_util import say
from gpusort_amd import datagen

pytestmark = pytest.mark.gpu

MASK32 = 0xFFFFFFFF


def u32_as_i64(t):
    return t.view(torch.int32).to(torch.int64) & MASK32


def test_single_rank_2p24_u32_equals_reference_merge_sort(handle):
    n = 1 << 24
    host_keys = datagen.uniform_u32(1, n)
    keys = torch.empty(n, dtype=torch.uint32, device="cuda")
    handle.generate(keys, g.GEN_UNIFORM_U32, 1)
    assert np.array_equal(keys.cpu().numpy(), host_keys), "device and host generators agree"
    say(f"single rank u32 n={n}")
    handle.sort_keys(keys)
    assert np.array_equal(keys.cpu().numpy(), oracle.merge_sort(host_keys))


def test_stable_kv_2p28_u32(handle):
    n = 1 << 28
    keys0 = torch.empty(n, dtype=torch.uint32, device="cuda")
    handle.generate(keys0, g.GEN_UNIFORM_U32, 1)
    vals = torch.empty(n, dtype=torch.uint32, device="cuda")
    handle.generate(vals, g.GEN_IOTA_U32, 0)
    keys = keys0.clone()
    say(f"stable KV u32+u32 n={n}")
    handle.sort_pairs(keys, vals)
    torch.cuda.synchronize()
    k = u32_as_i64(keys)
    idx = u32_as_i64(vals)
    assert bool((k[1:] >= k[:-1]).all()), "keys in order"
    assert bool((u32_as_i64(keys0)[idx] == k).all()), "payload maps input to output"
    assert bool(((idx[1:] > idx[:-1]) | (k[1:] != k[:-1])).all()), "ties keep input order"
    assert int(torch.bincount(idx, minlength=n).max()) == 1, "payloads are a permutation"


def test_vrank_hyksort_64_ranks_2p21_u64(handle):
    p, n = 64, 1 << 27
    off = np.arange(p + 1, dtype=np.uint64) * np.uint64(n // p)
    hk = datagen.uniform_u64(3, n)
    hv = np.arange(n, dtype=np.uint32)
    keys = torch.empty(n, dtype=torch.uint64, device="cuda")
    handle.generate(keys, g.GEN_UNIFORM_U64, 3)
    vals = torch.empty(n, dtype=torch.uint32, device="cuda")
    handle.generate(vals, g.GEN_IOTA_U32, 0)
    assert np.array_equal(keys.cpu().numpy(), hk), "device and host generators agree"
    say(f"vrank HykSort p={p} n={n} k=8 s=32")
    rep = handle.vrank_sort(keys, off, vals, algorithm=g.HYKSORT, k_way=8, samples_per_rank=32)
    wk, wv, want = oracle.vrank(hk, off.astype(np.int64), hv, algorithm="hyksort", k_way=8, samples_per_rank=32)
    assert np.array_equal(keys.cpu().numpy(), wk), "keys"
    assert np.array_equal(vals.cpu().numpy(), wv), "payloads"
    assert np.array_equal(rep["splitters"], want["splitters"]), "splitters"
    assert np.array_equal(rep["bucket_sizes"].astype(np.int64), want["bucket_sizes"]), "bucket sizes"
    assert np.array_equal(rep["final_rank_sizes"].astype(np.int64), want["final_rank_sizes"]), "final rank sizes"
    assert rep["splitter_iterations"] == want["splitter_iterations"]
    assert rep["max_rank_error"] <= 0.01 / 7


def test_vrank_64_ranks_tied_keys_random_payloads(handle):
    """The 64-rank HykSort shape at 2^19 keys per rank, with 1000 distinct
    u64 keys and random u32 payloads: every merge partition meets equal keys
    whose payloads decide the order (the merge's payload-order check and
    re-merge), checked against the oracle's re-run and the (key, payload)
    order of the input."""
    p, n = 64, 1 << 25
    off = np.arange(p + 1, dtype=np.uint64) * np.uint64(n // p)
    pool = datagen.uniform_u64(11, 1000)
    hk = pool[datagen.draws(12, n) % np.uint64(1000)]
    hv = (datagen.draws(13, n) >> np.uint64(32)).astype(np.uint32)
    keys = torch.from_numpy(hk.view(np.int64)).cuda().view(torch.uint64)
    vals = torch.from_numpy(hv.view(np.int32)).cuda().view(torch.uint32)
    say(f"vrank HykSort tied keys p={p} n={n} k=8 s=32")
    rep = handle.vrank_sort(keys, off, vals, algorithm=g.HYKSORT, k_way=8, samples_per_rank=32)
    wk, wv, want = oracle.vrank(hk, off.astype(np.int64), hv, algorithm="hyksort", k_way=8, samples_per_rank=32)
    gk, gv = keys.cpu().numpy(), vals.cpu().numpy()
    assert np.array_equal(gk, wk), "keys"
    assert np.array_equal(gv, wv), "payloads"
    assert np.array_equal(rep["splitters"], want["splitters"]), "splitters"
    assert np.array_equal(rep["bucket_sizes"].astype(np.int64), want["bucket_sizes"]), "bucket sizes"
    assert np.array_equal(rep["final_rank_sizes"].astype(np.int64), want["final_rank_sizes"]), "final rank sizes"
    order = np.lexsort((hv, hk))
    assert np.array_equal(gk, hk[order]) and np.array_equal(gv, hv[order]), "(key, payload) order"


def test_batched_100k_lognormal_arrays(handle):
    lengths = datagen.lognormal_lengths(4, 100_000, median=2048, sigma=2.0, lo=1, hi=1 << 22)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    n = int(off[-1])
    keys0 = torch.empty(n, dtype=torch.uint32, device="cuda")
    handle.generate(keys0, g.GEN_UNIFORM_U32, 4)
    starts = torch.from_numpy(off[:-1]).cuda()
    lens_t = torch.from_numpy(lengths.astype(np.int64)).cuda()
    seg_start = torch.repeat_interleave(starts, lens_t, output_size=n)
    vals = (torch.arange(n, dtype=torch.int64, device="cuda") - seg_start).to(torch.int32).view(torch.uint32)
    keys = keys0.clone()
    say(f"batched arrays=100000 n={n}")
    handle.sort_segments(keys, torch.from_numpy(off).cuda(), vals)
    torch.cuda.synchronize()
    k = u32_as_i64(keys)
    idx = u32_as_i64(vals)
    first = torch.zeros(n, dtype=torch.bool, device="cuda")
    first[starts] = True
    rest = ~first[1:]
    assert bool(((k[1:] >= k[:-1]) | ~rest).all()), "every array in order"
    src = seg_start + idx
    assert bool((idx < torch.repeat_interleave(lens_t, lens_t, output_size=n)).all()), "payload within its array"
    assert bool((u32_as_i64(keys0)[src] == k).all()), "payload maps input to output"
    assert bool((((idx[1:] > idx[:-1]) | (k[1:] != k[:-1])) | ~rest).all()), "ties keep input order"
    assert int(torch.bincount(src, minlength=n).max()) == 1, "payloads are a permutation of each array"


def test_served_requests_up_to_2p28_mixed_types(handle):
    count = 24
    sizes = datagen.loguniform_sizes(5, count, lo=1 << 10, hi=1 << 28)
    sizes[0] = 1 << 28  # the largest size of the range at least once
    kinds = [i % 3 for i in range(count)]  # u32, u64 + u32, f32
    bufs, expect = [], []
    for i, n in enumerate(int(x) for x in sizes):
        if kinds[i] == 0:
            k = torch.empty(n, dtype=torch.uint32, device="cuda")
            handle.generate(k, g.GEN_UNIFORM_U32, 1000 + i)
            v = None
        elif kinds[i] == 1:
            k = torch.empty(n, dtype=torch.uint64, device="cuda")
            handle.generate(k, g.GEN_UNIFORM_U64, 1000 + i)
            v = torch.empty(n, dtype=torch.uint32, device="cuda")
            handle.generate(v, g.GEN_IOTA_U32, 0)
        else:
            k = torch.empty(n, dtype=torch.float32, device="cuda")
            handle.generate(k, g.GEN_NORMAL_F32, 1000 + i, param=1e3)
            v = None
        ek, ev = k.clone(), None if v is None else v.clone()
        if ev is None:
            handle.sort_keys(ek)
        else:
            handle.sort_pairs(ek, ev)
        bufs.append((k, v))
        expect.append((ek, ev))
    torch.cuda.synchronize()
    say(f"served requests={count} largest={int(sizes.max())}")
    srv = g.Server(serial=False, streams=4, hbm_budget=200 * 10**9, coalesce_max_keys=1 << 16)
    ids = [srv.submit(k, v) for k, v in bufs]
    srv.drain()
    srv.close()

    def bits(t):
        return t.view(torch.int32) if t.element_size() == 4 else t.view(torch.int64)

    for i, ((k, v), (ek, ev)) in enumerate(zip(bufs, expect)):
        assert torch.equal(bits(k), bits(ek)), f"request {i} keys (n={k.numel()})"
        if v is not None:
            assert torch.equal(bits(v), bits(ev)), f"request {i} payloads"
    assert len(ids) == count
