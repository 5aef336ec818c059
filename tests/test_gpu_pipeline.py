"""Parity of the per-rank stages: merge path, bucket bounds, the virtual-rank
HykSort / SampleSort / HyperQuickSort pipeline, and the batched segmented sort.

The virtual-rank results (sorted data, splitters, bucket sizes, final rank
sizes, splitter quality) must equal the oracle's CPU re-run of the same rule
exactly."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from gpusort_amd import HYKSORT, HYPERQUICKSORT, SAMPLESORT, datagen
from tests.gpu_util import assert_bits_equal, dev, host, say

pytestmark = pytest.mark.gpu

ALGOS = {"hyksort": HYKSORT, "samplesort": SAMPLESORT, "hyperquicksort": HYPERQUICKSORT}


def keys_of(dtype, n, seed, distinct=None):
    d = np.dtype(dtype)
    r = datagen.draws(seed, n)
    if distinct:
        r = datagen.draws(seed + 1, distinct)[r % np.uint64(distinct)]
    if d == np.float32:
        return ((r >> np.uint64(40)).astype(np.float64) - 2.0**23).astype(np.float32)
    if d == np.float64:
        return (r >> np.uint64(12)).astype(np.float64) - 2.0**51
    if d.itemsize == 4:
        return (r >> np.uint64(32)).astype(np.uint32).view(d)
    return r.view(d)


@pytest.mark.parametrize("dtype", [np.uint32, np.int64, np.float32], ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("descending", [False, True], ids=["asc", "desc"])
def test_merge_keys_and_pairs(handle, dtype, descending):
    for na, nb in [(0, 5), (5, 0), (1, 1), (2047, 1), (2048, 2048), (3000, 70001), (100003, 99991)]:
        a = oracle.std_sort(keys_of(dtype, na, na + 3, distinct=max(1, na // 7)), descending)
        b = oracle.std_sort(keys_of(dtype, nb, nb + 5, distinct=max(1, nb // 7)), descending)
        say(f"merge {np.dtype(dtype).name} na={na} nb={nb} desc={descending}")
        out = dev(np.zeros(na + nb, dtype=dtype))
        handle.merge(dev(a), dev(b), out, descending=descending)
        assert_bits_equal(host(out), oracle.merge(a, b, descending), "merged keys")
        # pairs: payload shows which run and position every element came from
        va = np.arange(na, dtype=np.uint32)
        vb = np.arange(na, na + nb, dtype=np.uint32)
        ok, ov = dev(np.zeros(na + nb, dtype=dtype)), dev(np.zeros(na + nb, dtype=np.uint32))
        handle.merge(dev(a), dev(b), ok, dev(va), dev(vb), ov, descending=descending)
        cat = np.concatenate([a, b])
        want = np.argsort(-cat if descending else cat, kind="stable") if np.dtype(dtype).kind == "f" else \
            np.argsort(~cat if descending else cat, kind="stable")
        assert_bits_equal(host(ov), want.astype(np.uint32), "merge stability (ties from the first run)")


@pytest.mark.parametrize("upper", [False, True], ids=["lower", "upper"])
def test_bucket_bounds(handle, upper):
    runs = [oracle.std_sort(keys_of(np.uint64, m, m, distinct=50)) for m in [1, 7, 1000, 4096, 33333]]
    off = np.cumsum([0] + [r.size for r in runs]).astype(np.uint64)
    allk = np.concatenate(runs)
    spl = oracle.std_sort(keys_of(np.uint64, 9, 77, distinct=50))
    say(f"bucket bounds upper={upper}")
    got = handle.bucket_bounds(dev(allk), off, spl, upper=upper)
    for i, r in enumerate(runs):
        want = oracle.bounds(r, spl, upper=upper)
        assert got[i, 0] == off[i] and got[i, -1] == off[i + 1]
        assert np.array_equal(got[i, 1:-1] - off[i], want)


def rank_offsets(p, seed, lo, hi):
    sizes = lo + (datagen.draws(seed, p) % np.uint64(hi - lo + 1)).astype(np.int64)
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)


CASES = [
    # algorithm, p, k, samples, key dtype, payload, descending, distinct keys
    ("hyksort", 1, 8, 32, np.uint64, True, False, None),
    ("hyksort", 2, 2, 8, np.uint32, False, False, None),
    ("hyksort", 8, 4, 16, np.uint32, True, False, None),
    ("hyksort", 64, 8, 32, np.uint64, True, False, None),
    ("hyksort", 16, 8, 32, np.uint64, True, True, None),
    ("hyksort", 8, 2, 32, np.int64, False, True, 37),
    ("hyksort", 32, 4, 8, np.float32, True, False, None),
    ("hyksort", 16, 4, 32, np.uint32, True, False, 5),  # heavy ties: empty ranks rebalanced
    ("hyksort", 6, 8, 32, np.uint32, False, False, None),  # p not a power of 2 -> SampleSort
    ("hyksort", 32, 16, 8, np.uint64, True, False, None),  # 16 runs per destination: two merge passes
    ("hyksort", 64, 64, 4, np.int32, False, True, None),   # one round of 64 buckets: two merge passes
    ("hyksort", 16, 2, 1, np.uint32, True, False, None),   # one sample per rank: refined over iterations
    ("hyksort", 8, 4, 16, np.uint32, True, False, 1),      # all keys equal: selection ends on empty draws
    ("hyksort", 16, 8, 8, np.uint64, False, True, 1),
    ("samplesort", 8, 0, 0, np.uint32, True, False, None),
    ("samplesort", 64, 0, 0, np.uint64, True, False, None),
    ("samplesort", 12, 0, 0, np.int64, False, True, None),
    ("samplesort", 16, 0, 0, np.uint32, True, False, 3),
    ("hyperquicksort", 8, 0, 32, np.uint32, True, False, None),
    ("hyperquicksort", 64, 0, 32, np.uint64, False, False, None),
    ("hyperquicksort", 16, 0, 16, np.float32, True, True, None),
    ("hyperquicksort", 8, 0, 32, np.uint32, True, False, 4),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-p{c[1]}-k{c[2]}-{np.dtype(c[4]).name}"
                         f"{'-kv' if c[5] else ''}{'-desc' if c[6] else ''}{'-d' + str(c[7]) if c[7] else ''}")
def test_vrank_matches_cpu_rule(handle, case):
    algo, p, k, s, dtype, kv, desc, distinct = case
    off = rank_offsets(p, p * 31 + 7, 200 * p, 600 * p)
    n = int(off[-1])
    keys = keys_of(dtype, n, p + n, distinct=distinct)
    vals = np.arange(n, dtype=np.uint32) if kv else None
    say(f"vrank {case} n={n}")
    want_k, want_v, want = oracle.vrank(keys, off.astype(np.int64), vals, algorithm=algo, k_way=max(k, 2),
                                        samples_per_rank=max(s, 1), descending=desc)
    tk = dev(keys)
    tv = dev(vals) if kv else None
    got = handle.vrank_sort(tk, off, tv, algorithm=ALGOS[algo], k_way=max(k, 2), samples_per_rank=max(s, 1),
                            descending=desc)
    assert_bits_equal(host(tk), want_k, "keys")
    if kv:
        assert_bits_equal(host(tv), want_v, "payloads")
    assert got["rounds"] == want["rounds"]
    assert_bits_equal(got["splitters"], want["splitters"], "splitters")
    assert np.array_equal(got["bucket_sizes"].astype(np.int64), want["bucket_sizes"]), "bucket sizes"
    assert np.array_equal(got["final_rank_sizes"].astype(np.int64), want["final_rank_sizes"]), "final rank sizes"
    assert got["max_rank_error"] == pytest.approx(want["max_rank_error"], rel=1e-12, abs=1e-15)
    if algo == "hyksort" and p > 1 and (p & (p - 1)) == 0:
        assert got["splitter_iterations"] == want["splitter_iterations"]
    # the concatenation is the lexicographic sort of the input
    order = np.lexsort((vals, keys)) if kv else np.argsort(keys, kind="stable")
    if desc:
        order = order[::-1]
    assert_bits_equal(want_k, keys[order], "oracle keys are sorted")


LARGE_CASES = [
    # algorithm, p, k, samples, key dtype, payload, descending, distinct keys; 20K-60K keys per rank,
    # so every destination's merge spans many sampled partitions
    ("hyksort", 8, 8, 32, np.uint64, True, True, None),
    ("samplesort", 12, 0, 0, np.int32, True, False, 7),  # 12 runs per destination: two k-way passes
    ("hyperquicksort", 4, 0, 32, np.float64, False, False, None),
    ("hyksort", 16, 4, 32, np.uint32, True, False, 2),
]


@pytest.mark.parametrize("case", LARGE_CASES, ids=lambda c: f"{c[0]}-p{c[1]}-k{c[2]}-{np.dtype(c[4]).name}"
                         f"{'-kv' if c[5] else ''}{'-desc' if c[6] else ''}{'-d' + str(c[7]) if c[7] else ''}")
def test_vrank_large_ranks(handle, case):
    algo, p, k, s, dtype, kv, desc, distinct = case
    off = rank_offsets(p, p * 17 + 3, 20000, 60000)
    n = int(off[-1])
    keys = keys_of(dtype, n, p * 5 + n, distinct=distinct)
    vals = np.arange(n, dtype=np.uint32) if kv else None
    say(f"vrank large {case} n={n}")
    want_k, want_v, want = oracle.vrank(keys, off.astype(np.int64), vals, algorithm=algo, k_way=max(k, 2),
                                        samples_per_rank=max(s, 1), descending=desc)
    tk = dev(keys)
    tv = dev(vals) if kv else None
    got = handle.vrank_sort(tk, off, tv, algorithm=ALGOS[algo], k_way=max(k, 2), samples_per_rank=max(s, 1),
                            descending=desc)
    assert_bits_equal(host(tk), want_k, "keys")
    if kv:
        assert_bits_equal(host(tv), want_v, "payloads")
    assert np.array_equal(got["final_rank_sizes"].astype(np.int64), want["final_rank_sizes"]), "final rank sizes"
    assert_bits_equal(got["splitters"], want["splitters"], "splitters")
    assert np.array_equal(got["bucket_sizes"].astype(np.int64), want["bucket_sizes"]), "bucket sizes"
    if algo == "hyksort":
        assert got["splitter_iterations"] == want["splitter_iterations"]
        if distinct is None:
            assert got["max_rank_error"] <= 0.01 / (min(k, p) - 1)


TIE_CASES = [
    # algorithm, p, k, key dtype, payload dtype, descending, distinct keys, distinct payloads
    ("hyksort", 16, 4, np.uint32, np.uint32, False, 3, None),
    ("hyksort", 8, 8, np.uint64, np.uint32, True, 50, None),
    ("hyksort", 8, 4, np.uint32, np.uint64, False, 1, None),
    ("hyksort", 16, 2, np.uint64, np.int64, False, 5, 7),  # equal (key, payload) pairs
    ("samplesort", 12, 0, np.int32, np.uint32, False, 4, None),
    ("hyperquicksort", 8, 0, np.float32, np.uint64, True, 6, 3),
]


@pytest.mark.parametrize("case", TIE_CASES, ids=lambda c: f"{c[0]}-p{c[1]}-k{c[2]}-{np.dtype(c[3]).name}-"
                         f"{np.dtype(c[4]).name}{'-desc' if c[5] else ''}-d{c[6]}{'-v' + str(c[7]) if c[7] else ''}")
def test_vrank_random_payloads_on_tied_keys(handle, case):
    """Few distinct keys with payloads in no particular order: every merge
    meets equal keys whose payloads decide the order, across the pieces of
    one partition as well as within them."""
    algo, p, k, kdt, vdt, desc, distinct, vdistinct = case
    off = rank_offsets(p, p * 7 + 11, 3000, 9000)
    n = int(off[-1])
    keys = keys_of(kdt, n, 3 * p + n, distinct=distinct)
    r = datagen.draws(5 * n + p, n)
    if vdistinct:
        r = r % np.uint64(vdistinct)
    vals = (r >> np.uint64(32)).astype(np.uint32) if np.dtype(vdt).itemsize == 4 and not vdistinct else r.astype(vdt)
    say(f"vrank ties {case} n={n}")
    want_k, want_v, want = oracle.vrank(keys, off.astype(np.int64), vals, algorithm=algo, k_way=max(k, 2),
                                        samples_per_rank=16, descending=desc)
    tk, tv = dev(keys), dev(vals)
    got = handle.vrank_sort(tk, off, tv, algorithm=ALGOS[algo], k_way=max(k, 2), samples_per_rank=16,
                            descending=desc)
    assert_bits_equal(host(tk), want_k, "keys")
    assert_bits_equal(host(tv), want_v, "payloads")
    assert np.array_equal(got["final_rank_sizes"].astype(np.int64), want["final_rank_sizes"]), "final rank sizes"
    assert_bits_equal(got["splitters"], want["splitters"], "splitters")
    assert np.array_equal(got["bucket_sizes"].astype(np.int64), want["bucket_sizes"]), "bucket sizes"
    # the concatenation is the (key, payload) order of the input
    order = np.lexsort((vals, keys))
    if desc:
        order = order[::-1]
    assert_bits_equal(want_k, keys[order], "oracle keys in (key, payload) order")
    assert_bits_equal(want_v, vals[order], "oracle payloads in (key, payload) order")


SMALL_CASES = [
    # algorithm, p, key dtype, payload, descending, distinct keys: N < 5 p^2 takes the bitonic network
    ("samplesort", 2, np.uint32, True, False, None),
    ("samplesort", 3, np.uint32, False, False, None),
    ("samplesort", 5, np.uint64, True, True, None),
    ("samplesort", 6, np.int32, True, False, 7),
    ("samplesort", 8, np.float32, False, True, None),
    ("samplesort", 12, np.uint32, True, False, 3),
    ("samplesort", 13, np.int64, True, True, None),
    ("samplesort", 64, np.uint64, True, False, None),
    ("hyksort", 48, np.uint32, True, False, 50),  # p not a power of 2 -> SampleSort -> small path
]


@pytest.mark.parametrize("case", SMALL_CASES, ids=lambda c: f"{c[0]}-p{c[1]}-{np.dtype(c[2]).name}"
                         f"{'-kv' if c[3] else ''}{'-desc' if c[4] else ''}{'-d' + str(c[5]) if c[5] else ''}")
def test_vrank_small_input_bitonic_network(handle, case):
    algo, p, dtype, kv, desc, distinct = case
    off = rank_offsets(p, p * 13 + 1, 1, 4 * p)
    n = int(off[-1])
    assert n < 5 * p * p
    keys = keys_of(dtype, n, p + 3 * n, distinct=distinct)
    vals = np.arange(n, dtype=np.uint32) if kv else None
    say(f"small vrank {case} n={n}")
    want_k, want_v, want = oracle.vrank(keys, off.astype(np.int64), vals, algorithm=algo, k_way=8,
                                        samples_per_rank=32, descending=desc)
    tk = dev(keys)
    tv = dev(vals) if kv else None
    got = handle.vrank_sort(tk, off, tv, algorithm=ALGOS[algo], k_way=8, samples_per_rank=32, descending=desc)
    assert_bits_equal(host(tk), want_k, "keys")
    if kv:
        assert_bits_equal(host(tv), want_v, "payloads")
    assert got["rounds"] == want["rounds"] > 0
    assert np.array_equal(got["final_rank_sizes"].astype(np.int64), want["final_rank_sizes"]), "final rank sizes"
    assert len(got["splitters"]) == 0


def test_vrank_empty_rank_rejected(handle):
    from gpusort_amd import GpusortError

    off = np.array([0, 10, 10, 30], dtype=np.uint64)
    with pytest.raises(GpusortError) as e:
        handle.vrank_sort(dev(np.arange(30, dtype=np.uint32)), off)
    assert e.value.status == 7 and "No data to sort." in e.value.message


@pytest.mark.parametrize("kv", [False, True], ids=["keys", "pairs"])
def test_sort_segments(handle, kv):
    lengths = datagen.lognormal_lengths(3, 3000, median=300, sigma=2.0, lo=0, hi=1 << 17)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint64)
    n = int(off[-1])
    keys = keys_of(np.uint32, n, 9, distinct=1000)
    vals = np.arange(n, dtype=np.uint32)
    say(f"segments n={n} kv={kv}")
    tk, tv = dev(keys), dev(vals)
    handle.sort_segments(tk, dev(off), tv if kv else None)
    gk, gv = host(tk), host(tv)
    for s in range(lengths.size):
        a, b = int(off[s]), int(off[s + 1])
        perm = np.argsort(keys[a:b], kind="stable")
        assert np.array_equal(gk[a:b], keys[a:b][perm]), f"segment {s}"
        if kv:
            assert np.array_equal(gv[a:b], vals[a:b][perm]), f"segment {s} payload"


def _drawn_cases(seed, count):
    """Seeded configurations beyond the hand-picked ones: algorithm, ranks
    (powers of two and others), k, samples per rank, key type, payload,
    direction and number of distinct keys, all drawn from one stream."""
    d = datagen.draws(seed, 8 * count)
    algos = ["hyksort", "samplesort", "hyperquicksort"]
    dtypes = [np.uint32, np.int32, np.float32, np.uint64, np.int64, np.float64]
    out = []
    for c in range(count):
        x = [int(v) for v in d[8 * c: 8 * c + 8]]
        algo = algos[x[0] % 3]
        p = [1, 2, 3, 4, 5, 8, 12, 16, 32, 64][x[1] % 10]
        k = [2, 4, 8, 16][x[2] % 4] if algo == "hyksort" else 0
        s = [1, 4, 16, 32][x[3] % 4] if algo != "samplesort" else 0
        distinct = [None, None, 3, 50, 1000][x[4] % 5]
        out.append((algo, p, k, s, dtypes[x[5] % 6], bool(x[6] & 1), bool(x[7] & 1), distinct))
    return out


@pytest.mark.parametrize("case", _drawn_cases(2024, 16),
                         ids=lambda c: f"{c[0]}-p{c[1]}-k{c[2]}-s{c[3]}-{np.dtype(c[4]).name}"
                         f"{'-kv' if c[5] else ''}{'-desc' if c[6] else ''}{'-d' + str(c[7]) if c[7] else ''}")
def test_vrank_drawn_configurations(handle, case):
    algo, p, k, s, dtype, kv, desc, distinct = case
    off = rank_offsets(p, p * 17 + 3, 150 * p, 450 * p)
    n = int(off[-1])
    keys = keys_of(dtype, n, 5 * p + n, distinct=distinct)
    vals = np.arange(n, dtype=np.uint32) if kv else None
    say(f"vrank drawn {case} n={n}")
    want_k, want_v, want = oracle.vrank(keys, off.astype(np.int64), vals, algorithm=algo, k_way=max(k, 2),
                                        samples_per_rank=max(s, 1), descending=desc)
    tk = dev(keys)
    tv = dev(vals) if kv else None
    got = handle.vrank_sort(tk, off, tv, algorithm=ALGOS[algo], k_way=max(k, 2), samples_per_rank=max(s, 1),
                            descending=desc)
    assert_bits_equal(host(tk), want_k, "keys")
    if kv:
        assert_bits_equal(host(tv), want_v, "payloads")
    assert got["rounds"] == want["rounds"]
    assert_bits_equal(got["splitters"], want["splitters"], "splitters")
    assert np.array_equal(got["bucket_sizes"].astype(np.int64), want["bucket_sizes"]), "bucket sizes"
    assert np.array_equal(got["final_rank_sizes"].astype(np.int64), want["final_rank_sizes"]), "final rank sizes"
