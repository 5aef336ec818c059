"""Pins the CPU oracle (oracle/oracle.cpp) before its results are trusted in the parity tests.

* Golden vectors of the reference's own OmpUtils (tests/golden/, produced by
  make_golden.py from include/omp/omp_utils.h compiled out of the reference
  tree): the restatement must reproduce them bit for bit.
* The reference's compiled OmpUtils itself (oracle/_ref), when built here.
* Closed forms the reference's tests use as known answers
  (src/test/arithmetic_progression_checker.h:49-106, ln_checker.h:45-87,
  log_checker.h:45-87, zero_checker.h:34-74) and library routines (sort,
  stable argsort, lexsort, searchsorted, cumsum).
* Brute force over every permutation of tiny inputs, with thread counts that
  take the split-and-merge path of MergeSort.
"""
from __future__ import annotations

import itertools
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from gpusort_amd import datagen

GOLDEN = Path(__file__).resolve().parent / "golden" / "omp_utils_golden.npz"


@pytest.fixture(autouse=True)
def restore_threads():
    yield
    oracle.set_threads(0)


def _golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_golden_merge_sort(threads):
    g = _golden()
    oracle.set_threads(threads)
    tags = sorted({k[: -len("_in")] for k in g.files if k.startswith(f"sort_t{threads}_") and k.endswith("_in")})
    assert tags
    for tag in tags:
        desc = tag.endswith("_d1")
        got = oracle.merge_sort(g[tag + "_in"], descending=desc)
        want = g[tag + "_out"]
        assert got.dtype == want.dtype and np.array_equal(got.view(np.uint8), want.view(np.uint8)), tag


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_golden_pairs_merge(threads):
    g = _golden()
    oracle.set_threads(threads)
    for tag in sorted({k[: -len("_kin")] for k in g.files if k.startswith(f"pairs_t{threads}_") and k.endswith("_kin")}):
        k, v = oracle.merge_sort_pairs(g[tag + "_kin"], g[tag + "_vin"], descending=tag.endswith("_d1"))
        assert np.array_equal(k, g[tag + "_kout"]) and np.array_equal(v, g[tag + "_vout"]), tag
    for tag in sorted({k[: -len("_out")] for k in g.files if k.startswith(f"merge_t{threads}_") and k.endswith("_out")}):
        assert np.array_equal(oracle.merge(g[tag + "_a"], g[tag + "_b"]), g[tag + "_out"]), tag
    for tag in sorted({k[: -len("_out")] for k in g.files if k.startswith(f"scan_t{threads}_") and k.endswith("_out")}):
        assert np.array_equal(oracle.scan(g[tag + "_a"], first=3), g[tag + "_out"]), tag


def test_reference_lib_agrees_when_built():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for threads in (2, 5, 8):
        oracle.set_threads(threads)
        for dtype in (np.uint32, np.float64, np.int64):
            k = datagen.uniform_u64(threads, 50_003).view(np.int64).astype(dtype) if dtype != np.uint32 else \
                datagen.uniform_u32(threads, 50_003)
            assert np.array_equal(oracle.merge_sort(k), oracle.ref_merge_sort(k))
            assert np.array_equal(oracle.merge_sort(k, True), oracle.ref_merge_sort(k, True))
        k = datagen.uniform_u32(9, 40_000) % np.uint32(300)
        v = datagen.uniform_u32(10, 40_000) % np.uint32(7)
        for desc in (False, True):
            a = oracle.merge_sort_pairs(k, v, desc)
            b = oracle.ref_merge_sort_pairs(k, v, desc)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("threads", [1, 4, 8])
@pytest.mark.parametrize("dtype", [np.uint32, np.int32, np.float32, np.uint64, np.int64, np.float64],
                         ids=lambda d: np.dtype(d).name)
def test_merge_sort_equals_sort(threads, dtype):
    oracle.set_threads(threads)
    r = datagen.draws(threads * 7 + np.dtype(dtype).itemsize, 33_333)
    k = (r >> np.uint64(16)).astype(np.float64).astype(dtype) if np.dtype(dtype).kind == "f" else r.view(np.int64).astype(dtype)
    assert np.array_equal(oracle.merge_sort(k), np.sort(k, kind="stable"))
    assert np.array_equal(oracle.merge_sort(k, descending=True), np.sort(k)[::-1])


def test_arithmetic_progression_known_answer():
    # ArithmeticProgressionChecker: the sorted output is exactly a1 + i*d,
    # descending when d < 0 (src/test/arithmetic_progression_checker.h:49-106).
    oracle.set_threads(6)
    n = 100_000
    for a1, d in ((-5000, 3), (7, -2)):
        ap = (a1 + np.arange(n, dtype=np.int64) * d)
        shuffled = ap[np.argsort(datagen.uniform_u64(31, n), kind="stable")]
        assert np.array_equal(oracle.merge_sort(shuffled, descending=d < 0), ap)


def test_log_sequences_known_answer():
    # LnChecker / LogChecker: values log(i) / log10(i) of the global index
    # (src/test/ln_checker.h:45-87, log_checker.h:45-87).
    n = 50_000
    idx = np.arange(1, n + 1, dtype=np.float64)
    for f in (np.log, np.log10):
        seq = f(idx)
        shuffled = seq[np.argsort(datagen.uniform_u64(5, n), kind="stable")]
        got = oracle.merge_sort(shuffled)
        assert np.array_equal(got, seq)
        assert got[0] == 0.0 and got[-1] == f(float(n))


def test_all_equal_known_answer():
    # ZeroChecker: every element equals the one constant (zero_checker.h:34-74).
    k = np.full(77_777, 42, dtype=np.int32)
    assert np.array_equal(oracle.merge_sort(k), k)
    kk, vv = oracle.stable_sort_pairs(k, np.arange(k.size, dtype=np.uint32))
    assert np.array_equal(vv, np.arange(k.size, dtype=np.uint32))


@pytest.mark.parametrize("threads", [2, 3])
def test_brute_force_permutations(threads):
    # n >= 2p takes the split + merge path; every permutation of a multiset
    # with ties must come out sorted, and pairs in lexicographic order.
    oracle.set_threads(threads)
    base = np.array([3, 1, 3, 2, 1, 0, 3], dtype=np.int32)
    for perm in itertools.permutations(range(base.size)):
        k = base[list(perm)]
        assert np.array_equal(oracle.merge_sort(k), np.sort(base))
    vals = np.arange(7, dtype=np.uint32)
    for perm in itertools.islice(itertools.permutations(range(7)), 0, 5040, 7):
        k = base[list(perm)]
        v = vals[list(perm)]
        gk, gv = oracle.merge_sort_pairs(k, v)
        order = np.lexsort((v, k))
        assert np.array_equal(gk, k[order]) and np.array_equal(gv, v[order])
        sk, sv = oracle.stable_sort_pairs(k, v)
        st = np.argsort(k, kind="stable")
        assert np.array_equal(sv, v[st])


def test_merge_ties_take_first_run():
    # std::merge: equal elements of the first range precede those of the second.
    a = np.array([1, 2, 2, 5], dtype=np.int64)
    b = np.array([2, 2, 3], dtype=np.int64)
    assert np.array_equal(oracle.merge(a, b), np.array([1, 2, 2, 2, 2, 3, 5]))
    for threads in (1, 2, 4):
        oracle.set_threads(threads)
        x = np.sort(datagen.uniform_u32(3, 20_000) % np.uint32(500))
        y = np.sort(datagen.uniform_u32(4, 31_000) % np.uint32(500))
        assert np.array_equal(oracle.merge(x, y), np.sort(np.concatenate([x, y])))


def test_scan_closed_form():
    for cnt in (1, 50, 1000, 123_457):
        oracle.set_threads(7)
        a = (datagen.draws(cnt, cnt) % np.uint64(1000)).astype(np.int64)
        want = np.concatenate([[11], 11 + np.cumsum(a)[:-1]]) if cnt else a
        assert np.array_equal(oracle.scan(a, first=11), want)


def test_bounds_match_searchsorted():
    run = np.sort(datagen.uniform_u32(2, 10_000) % np.uint32(2000))
    spl = np.array([0, 1, 999, 1000, 1999, 2000, 4000], dtype=np.uint32)
    assert np.array_equal(oracle.bounds(run, spl), np.searchsorted(run, spl, side="left"))
    assert np.array_equal(oracle.bounds(run, spl, upper=True), np.searchsorted(run, spl, side="right"))
    rdesc = run[::-1].copy()
    neg = -rdesc.astype(np.int64)
    assert np.array_equal(oracle.bounds(rdesc, spl, descending=True),
                          np.searchsorted(neg, -spl.astype(np.int64), side="left"))


def test_keep_no_empty_rules():
    # Hand-run of ParUtils::KeepNoProcessHasEmptyData (par_utils.h:103-175):
    # nelems is the gathered vector (not updated in the loop) and the rank
    # after a handled partner is skipped.
    assert oracle.keep_no_empty([4, 5, 6]).tolist() == [4, 5, 6]
    assert oracle.keep_no_empty([0, 10, 7]).tolist() == [5, 5, 7]
    assert oracle.keep_no_empty([0, 0, 9, 1]).tolist() == [3, 3, 3, 1]
    assert oracle.keep_no_empty([4, 5, 0]).tolist() == [4, 3, 2]
    assert oracle.keep_no_empty([5, 0, 0, 9]).tolist() == [5, 3, 3, 3]
    assert oracle.keep_no_empty([0, 10, 0, 10]).tolist() == [5, 5, 0, 10]
    assert oracle.keep_no_empty([8, 0, 0]).tolist() == [4, 2, 2]
    s = datagen.uniform_u32(1, 64) % np.uint32(3)
    out = oracle.keep_no_empty(s.astype(np.int64))
    assert out.sum() == s.sum()


@pytest.mark.parametrize("algorithm", ["hyksort", "samplesort", "hyperquicksort"])
def test_vrank_rule_invariants(algorithm):
    p = 16
    sizes = 500 + (datagen.draws(4, p) % np.uint64(2000)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)])
    n = int(off[-1])
    k = datagen.uniform_u64(8, n) % np.uint64(n // 3)
    v = np.arange(n, dtype=np.uint32)
    gk, gv, rep = oracle.vrank(k, off, v, algorithm=algorithm, k_way=4, samples_per_rank=16)
    order = np.lexsort((v, k))
    assert np.array_equal(gk, k[order]) and np.array_equal(gv, v[order])
    assert rep["final_rank_sizes"].sum() == n
    b = rep["bucket_sizes"]
    per_round = b.size // max(rep["rounds"], 1)
    for r in range(rep["rounds"]):
        assert b[r * per_round:(r + 1) * per_round].sum() == n
    if algorithm == "hyksort":
        assert rep["rounds"] == 2 and rep["splitters"].size == 3 + 4 * 3
        assert 0 <= rep["max_rank_error"] < 0.25
    # final ranks are contiguous pieces of the global order
    fo = np.concatenate([[0], np.cumsum(rep["final_rank_sizes"])])
    assert fo[-1] == n


def test_vrank_tiny_hand_case():
    # p=2, k=2, one sample per rank: rank 0 = [5 1 9 3] -> sorted [1 3 5 9],
    # sample run[(0+1)*4/2] = 5; rank 1 = [8 2 6 4] -> [2 4 6 8], sample 6.
    # Sorted samples [5 6]; splitter = S[1*2/2 - 1] = 5.  lower_bound(5):
    # rank 0 -> [1 3 | 5 9], rank 1 -> [2 4 | 6 8].  Rank 0 receives the low
    # buckets, rank 1 the high ones.
    k = np.array([5, 1, 9, 3, 8, 2, 6, 4], dtype=np.uint32)
    gk, _, rep = oracle.vrank(k, [0, 4, 8], algorithm="hyksort", k_way=2, samples_per_rank=1)
    assert gk.tolist() == [1, 2, 3, 4, 5, 6, 8, 9]
    assert rep["splitters"].tolist() == [5]
    assert rep["bucket_sizes"].tolist() == [2, 2, 2, 2]
    assert rep["final_rank_sizes"].tolist() == [4, 4]
    assert rep["max_rank_error"] == pytest.approx(0.0)


def test_datagen_streams_are_counter_based():
    a = datagen.draws(5, 1000)
    b = datagen.draws(5, 400, start=600)
    assert np.array_equal(a[600:], b)
    # splitmix64 of state seed + G: the first output for seed 0 is the
    # published first value of the splitmix64 sequence seeded with 0.
    assert int(datagen.draws(0, 1)[0]) == 0xE220A8397B1DCDAF
    u = datagen.uniform_u32(1, 1 << 16)
    assert abs(float(u.mean()) / 2**32 - 0.5) < 0.01
    s = datagen.sorted_u32(3, 100_000)
    assert np.all(s[1:] >= s[:-1])
    z = datagen.zipf_u32(2, 200_000)
    top = np.unique(z, return_counts=True)[1].max() / z.size
    h = sum(1.0 / k**1.1 for k in range(1, 1 << 20))
    assert abs(top - 1.0 / h) < 0.01
    f = datagen.normal_f32(4, 200_000)
    assert np.all(np.isfinite(f)) and not np.any(np.signbit(f) & (f == 0))
    assert abs(float(f.std()) / 1e3 - 1.0) < 0.01
    L = datagen.lognormal_lengths(1, 100_000)
    assert L.min() >= 1 and L.max() <= 1 << 22
    assert abs(math.log(np.median(L)) - math.log(2048)) < 0.05


@pytest.mark.parametrize("descending", [False, True])
def test_segmented_oracle_matches_lexsort(descending):
    # Batched oracle = numpy's stable lexsort by (segment, key) (closed form).
    rng = np.random.default_rng(7)
    lengths = rng.integers(0, 300, 200)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    n = int(off[-1])
    k = rng.integers(-50, 50, n).astype(np.int64)
    v = np.arange(n, dtype=np.uint32)
    seg = np.repeat(np.arange(lengths.size), lengths)
    order = np.lexsort((-k if descending else k, seg))
    kk, vv = oracle.seg_stable_sort_pairs(k, v, off, descending)
    assert np.array_equal(kk, k[order]) and np.array_equal(vv, v[order])
    assert np.array_equal(oracle.seg_sort(k, off, descending), k[order])


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("descending", [False, True])
def test_multisplit_oracle_matches_searchsorted(upper, descending):
    # Bucket = searchsorted side (closed form); stable partition = stable argsort by bucket.
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 60, 5000).astype(np.int64)
    spl = np.sort(rng.choice(keys, 9, replace=False))
    if descending:
        spl = spl[::-1].copy()
        b = np.searchsorted(-spl, -keys, side="left" if upper else "right")
    else:
        b = np.searchsorted(spl, keys, side="left" if upper else "right")
    idx, offs = oracle.multisplit(keys, spl, upper, descending)
    assert np.array_equal(idx, np.argsort(b, kind="stable"))
    assert np.array_equal(offs, np.concatenate([[0], np.cumsum(np.bincount(b, minlength=spl.size + 1))]))


def test_merge_lists_known_answers():
    # Traced by hand through MergeLists (samplesort_helper.h:305-387).
    a, b = np.arange(1, 9), np.array([9, 10])
    assert oracle.merge_lists(a, b, True, True).tolist() == list(range(1, 9))      # low = 9 > high = 8
    assert oracle.merge_lists(b, a, False, True).tolist() == [9, 10]
    assert oracle.merge_lists([5, 5], [5, 5], True, True).tolist() == [5, 5]
    assert oracle.merge_lists([5, 5, 5], [5], False, True).tolist() == [5, 5]
    # decreasing lists: keep low keeps the smallest values (the tail)
    assert oracle.merge_lists([9, 7, 5], [8, 6, 4], True, False).tolist() == [6, 5, 4]
    assert oracle.merge_lists([8, 6, 4], [9, 7, 5], False, False).tolist() == [9, 8, 7]


def _merge_lists_closed_form(a, b, keep_low, inc):
    # Each loop condition of MergeLists holds on a prefix (or suffix) of the
    # sorted merge, so the kept length is a min/max of element counts.
    s = np.sort(np.concatenate([a, b]))
    if not inc:
        s = s[::-1]
    n, half = s.size, s.size // 2
    if inc:
        low, high = max(a[0], b[0]), min(a[-1], b[-1])
        if keep_low:
            L = min(max(int((s < low).sum()), half), int((s <= high).sum()))
            return s[:L]
        L = max(min(n - half, int((s >= low).sum())), int((s > high).sum()))
        return s[n - L:]
    low, high = max(a[-1], b[-1]), min(a[0], b[0])
    if keep_low:
        L = max(min(int((s <= high).sum()), n - half), int((s < low).sum()))
        return s[n - L:]
    L = max(min(half, int((s >= low).sum())), int((s > high).sum()))
    return s[:L]


def test_merge_lists_equals_closed_form():
    rng = np.random.default_rng(5)
    for trial in range(3000):
        na, nb = rng.integers(1, 12, 2)
        hi = int(rng.integers(2, 30))
        a = np.sort(rng.integers(0, hi, na))
        b = np.sort(rng.integers(0, hi, nb))
        for inc in (True, False):
            aa, bb = (a, b) if inc else (a[::-1].copy(), b[::-1].copy())
            for keep_low in (True, False):
                got = oracle.merge_lists(aa, bb, keep_low, inc)
                want = _merge_lists_closed_form(aa, bb, keep_low, inc)
                assert np.array_equal(got, want), (trial, aa, bb, keep_low, inc)


@pytest.mark.parametrize("p", [2, 3, 5, 6, 7, 8, 12, 13])
def test_small_samplesort_network_invariants(p):
    # N < 5 p^2: globally sorted, every element kept, no rank empty.
    rng = np.random.default_rng(p)
    sizes = rng.integers(1, 4 * p, p)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(off[-1])
    assert n < 5 * p * p
    k = rng.integers(0, 40, n).astype(np.uint32)
    v = np.arange(n, dtype=np.uint32)
    for desc in (False, True):
        ok, ov, rep = oracle.vrank(k, off, v, algorithm="samplesort", descending=desc)
        pairs = sorted(zip(k.tolist(), v.tolist()), reverse=desc)
        assert list(zip(ok.tolist(), ov.tolist())) == pairs
        assert (rep["final_rank_sizes"] > 0).all() and rep["final_rank_sizes"].sum() == n
        assert rep["rounds"] > 0 and len(rep["splitters"]) == 0


# ------------------------------------------------------------------ splitter-rule pins
# The oracle's SampleSort splitters are a closed form over the sorted samples.
# The reference does not sort the samples that way: each rank's p-1 samples
# arr[(i+1)*n/p] (samplesort_helper.h:402-405) go through the distributed
# bitonic network (BitonicSort, :192-235, :408), then every rank contributes
# send_splits[npes-2] (ascending) or send_splits[0] (descending) to an
# Allgather that drops the last (ascending) or first (descending) entry
# (:411-416).  The oracle's literal network (its N < 5p^2 path) runs that
# sort here on the sample lists alone, so the check goes through different
# code than the closed form and catches a wrong sample position or pick index.
def _local_runs(k, off, v, descending):
    runs = []
    for r in range(off.size - 1):
        a, b = int(off[r]), int(off[r + 1])
        if v is None:
            s = np.sort(k[a:b])
            runs.append((s[::-1].copy() if descending else s, None))
        else:
            order = np.lexsort((v[a:b], k[a:b]))
            if descending:
                order = order[::-1]
            runs.append((k[a:b][order], v[a:b][order]))
    return runs


@pytest.mark.parametrize("p", [2, 3, 5, 8, 12, 13, 64])
@pytest.mark.parametrize("descending", [False, True], ids=["asc", "desc"])
@pytest.mark.parametrize("with_vals", [False, True], ids=["keys", "pairs"])
def test_samplesort_splitters_through_the_reference_network(p, descending, with_vals):
    rng = np.random.default_rng(1000 + p)
    sizes = rng.integers(5 * p, 9 * p, p)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(off[-1])
    assert n >= 5 * p * p  # the sampling path, not the small-input network
    k = rng.integers(0, 3 * p, n).astype(np.uint32)  # many ties
    v = rng.permutation(n).astype(np.uint32) if with_vals else None
    _, _, rep = oracle.vrank(k, off, v, algorithm="samplesort", descending=descending)
    # Samples of every rank's sorted run, p-1 each.
    sk, sv = [], []
    for rk, rv in _local_runs(k, off, v, descending):
        m = rk.size
        idx = [(i + 1) * m // p for i in range(p - 1)]
        sk.append(rk[idx])
        if with_vals:
            sv.append(rv[idx])
    soff = np.arange(p + 1, dtype=np.int64) * (p - 1)
    sk_all = np.concatenate(sk)
    sv_all = np.concatenate(sv) if with_vals else None
    assert sk_all.size < 5 * p * p  # the network path
    nk, _, nrep = oracle.vrank(sk_all, soff, sv_all, algorithm="samplesort", descending=descending)
    # Equal lists stay equal under the network's merge-splits, so every rank
    # still holds p-1 samples.
    assert nrep["final_rank_sizes"].tolist() == [p - 1] * p
    blocks = nk.reshape(p, p - 1)
    picked = blocks[:, p - 2] if not descending else blocks[:, 0]
    want = picked[:-1] if not descending else picked[1:]
    assert np.array_equal(rep["splitters"], want), (rep["splitters"], want)


def test_samplesort_network_pick_catches_index_errors():
    # The comparison above is sensitive: picking the neighbouring sample or
    # dropping the other end gives different splitters on this input.
    p = 8
    rng = np.random.default_rng(3)
    sizes = rng.integers(5 * p, 9 * p, p)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    k = rng.integers(0, 10_000, int(off[-1])).astype(np.uint32)
    _, _, rep = oracle.vrank(k, off, algorithm="samplesort")
    s = np.sort(np.concatenate([np.sort(k[off[r]:off[r + 1]])[[(i + 1) * int(sizes[r]) // p for i in range(p - 1)]]
                                for r in range(p)]))
    right = s[[(r + 1) * (p - 1) - 1 for r in range(p - 1)]]
    assert np.array_equal(rep["splitters"], right)
    assert not np.array_equal(rep["splitters"], s[[(r + 1) * (p - 1) for r in range(p - 1)]])
    assert not np.array_equal(rep["splitters"], s[[(r + 1) * (p - 1) - 2 for r in range(p - 1)]])


def test_hyperquicksort_hand_traced_p2():
    # p = 2, two regular samples per rank at (j+1)*8/3 = positions 2 and 5.
    # rank 0 sorted [1 2 4 7 8 10 13 15] -> samples 4, 10
    # rank 1 sorted [0 3 5 6 9 11 12 14] -> samples 5, 11
    # gathered in rank order [4 10 5 11]; global lower_bound ranks 4 10 5 11;
    # target 16/2 = 8: distances 4 2 3 3 -> pivot 10 (hyperquicksort_helper.h:251-256).
    # lower_bound(10): rank 0 keeps [1 2 4 7 8] (5 low, 3 high), rank 1 sends
    # its low 5 and keeps [11 12 14].
    k = np.array([13, 2, 7, 10, 4, 15, 1, 8, 6, 12, 3, 14, 9, 0, 11, 5], dtype=np.int32)
    gk, _, rep = oracle.vrank(k, [0, 8, 16], algorithm="hyperquicksort", samples_per_rank=2)
    assert gk.tolist() == list(range(16))
    assert rep["splitters"].tolist() == [10]
    assert rep["bucket_sizes"].tolist() == [5, 3, 5, 3]
    assert rep["final_rank_sizes"].tolist() == [10, 6]
    assert rep["rounds"] == 1


def test_hyperquicksort_hand_traced_p4():
    # p = 4, two samples per rank at (j+1)*len/3, ragged ranks.
    # Round 1 (one group, 16 keys, target 8):
    #   r0 [3 8 11 17 20] samples 8 17; r1 [1 6 14] samples 6 14;
    #   r2 [0 5 9 12 15 19] samples 9 15; r3 [2 18] samples 2 18.
    #   rank order [8 17 6 14 9 15 2 18], global ranks 6 12 5 10 7 11 2 13,
    #   distances 2 4 3 2 1 3 6 5 -> pivot 9.  lower_bound(9): 2 2 2 1.
    #   Ranks 0,1 (<= (4-1)/2) keep their low sides; partners 0<->2, 1<->3
    #   (hyperquicksort_helper.h:358-372): r0 [0 3 5 8], r1 [1 2 6],
    #   r2 [9 11 12 15 17 19 20], r3 [14 18].
    # Round 2, group {0,1} (7 keys, target 3): samples r0 3 5, r1 2 6; rank
    #   of 3 is 3 -> pivot 3; lower_bound 1 and 2; r0 [0 1 2], r1 [3 5 6 8].
    # Round 2, group {2,3} (9 keys, target 4): samples r2 12 17, r3 14 18;
    #   ranks 2 5 3 6, distances 2 1 1 2 -> first best 17; lower_bound 4 and 1;
    #   r2 [9 11 12 14 15], r3 [17 18 19 20].
    k = np.array([20, 3, 17, 8, 11, 1, 14, 6, 9, 19, 0, 12, 5, 15, 2, 18], dtype=np.uint64)
    gk, _, rep = oracle.vrank(k, [0, 5, 8, 14, 16], algorithm="hyperquicksort", samples_per_rank=2)
    assert gk.tolist() == [0, 1, 2, 3, 5, 6, 8, 9, 11, 12, 14, 15, 17, 18, 19, 20]
    assert rep["splitters"].tolist() == [9, 3, 17]
    assert rep["bucket_sizes"].tolist() == [2, 3, 2, 1, 2, 4, 1, 1, 1, 3, 2, 1, 4, 3, 1, 1]
    assert rep["final_rank_sizes"].tolist() == [3, 4, 5, 4]
    assert rep["rounds"] == 2


def _exchange_rounds_bucket_sizes(k, off, algorithm, k_way, rep, descending=False):
    # Per (round, source rank, bucket) counts from np.searchsorted on the runs
    # each round starts from, with the round's reported splitters and the
    # reference's exchange structure (hyksort_helper.h:546-633,
    # samplesort_helper.h:504-565, hyperquicksort_helper.h:348-382).
    p = off.size - 1
    runs = [np.sort(k[off[r]:off[r + 1]]) for r in range(p)]
    if descending:
        runs = [r[::-1].copy() for r in runs]
    key = (lambda a: -a.astype(np.float64)) if descending else (lambda a: a.astype(np.float64))
    spl = rep["splitters"].tolist()
    out = []
    if algorithm == "samplesort":
        cuts = [np.concatenate([[0], np.searchsorted(key(r), key(np.array(spl, dtype=k.dtype)), side="right"),
                                [r.size]]) for r in runs]
        for c in cuts:
            out += np.diff(c).tolist()
        return out
    n_p, si = p, 0
    while n_p > 1:
        kk = min(k_way, n_p) if algorithm == "hyksort" else 2
        blk = n_p // kk
        nxt = [None] * p
        for g0 in range(0, p, n_p):
            s = np.array(spl[si:si + kk - 1], dtype=k.dtype)
            si += kk - 1
            cuts = [np.concatenate([[0], np.searchsorted(key(runs[r]), key(s), side="left"), [runs[r].size]])
                    for r in range(g0, g0 + n_p)]
            for c in cuts:
                out += np.diff(c).tolist()
            for dl in range(n_p):
                if algorithm == "hyksort":
                    i = dl // blk
                    parts = [runs[g0 + lr][cuts[lr][i]:cuts[lr][i + 1]] for lr in range(dl % blk, n_p, blk)]
                else:
                    low = dl <= (n_p - 1) // 2
                    partner = dl + (n_p // 2 if low else -(n_p // 2))
                    parts = [runs[g0 + q][cuts[q][0 if low else 1]:cuts[q][1 if low else 2]] for q in (dl, partner)]
                m = np.sort(np.concatenate(parts))
                nxt[g0 + dl] = m[::-1].copy() if descending else m
        runs = nxt
        n_p = blk
    return out


@pytest.mark.parametrize("algorithm,k_way", [("hyksort", 4), ("hyksort", 2), ("samplesort", 0),
                                             ("hyperquicksort", 0)])
@pytest.mark.parametrize("descending", [False, True], ids=["asc", "desc"])
def test_bucket_sizes_equal_searchsorted_counts(algorithm, k_way, descending):
    p = 16
    rng = np.random.default_rng(29)
    sizes = rng.integers(200, 900, p)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    k = rng.integers(0, 5000, int(off[-1])).astype(np.int64)
    _, _, rep = oracle.vrank(k, off, algorithm=algorithm, k_way=max(k_way, 2), samples_per_rank=8,
                             descending=descending)
    want = _exchange_rounds_bucket_sizes(k, off, algorithm, k_way, rep, descending)
    assert rep["bucket_sizes"].tolist() == want


def test_hyksort_selection_hand_traced():
    # p = 2, k = 2 (one splitter): tol = 0.01, beta = (int)(cbrt(100) * 3) = 13,
    # N = 40, target exp = 20, loop while max_err > (int)(40 * 0.01) = 0.
    # rank 0 = 0..15, rank 1 = 16..39 (sorted runs; value = global rank).
    # Iteration 1: one regular sample per rank, run[n/2]: 8 and 28; global
    #   ranks 8 and 28 -> lb = 8 (error 12), ub = 28.  Windows become
    #   [lower_bound(8), lower_bound(28)): rank 0 [8, 16), rank 1 [0, 12).
    # Iteration 2: tot = 8 + 12 = 20, c = min(13, 20) = 13 draws at
    #   u_j = (2j+1)*20/26 = 0 2 3 5 6 8 10 11 13 14 16 17 19 of the windows
    #   laid end to end: rank 0 values 8 10 11 13 14, rank 1 values
    #   16 18 19 21 22 24 25 27.  20 is not drawn: lb = 19 (error 1), ub = 21;
    #   windows rank 0 [16, 16), rank 1 [3, 5).
    # Iteration 3: tot = 2, draws u = 0, 1 -> values 19, 20 -> lb = 20, error 0.
    # (Draws at (j+1)*tot/(c+1) would have hit 20 in iteration 2.)
    k = np.arange(40, dtype=np.uint32)
    gk, _, rep = oracle.vrank(k, [0, 16, 40], algorithm="hyksort", k_way=2, samples_per_rank=1)
    assert gk.tolist() == list(range(40))
    assert rep["splitters"].tolist() == [20]
    assert rep["splitter_iterations"] == 3
    assert rep["bucket_sizes"].tolist() == [16, 0, 4, 20]
    assert rep["final_rank_sizes"].tolist() == [20, 20]
    assert rep["max_rank_error"] == 0.0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_hyksort_selection_reaches_tolerance(seed):
    # Uniform keys: the refined splitters land within tol = 0.01 / (k-1) of
    # their targets in every round (hyksort_helper.h:340-356), where the one
    # regular-sample pick alone misses by about 2.6 % of N at k = 8.
    p, n, kw = 64, 1 << 20, 8
    off = np.arange(p + 1, dtype=np.int64) * (n // p)
    k = datagen.uniform_u64(seed, n)
    _, _, rep = oracle.vrank(k, off, np.arange(n, dtype=np.uint32), algorithm="hyksort", k_way=kw,
                             samples_per_rank=32)
    assert rep["rounds"] == 2
    assert rep["max_rank_error"] <= 0.01 / (kw - 1)
    assert 2 <= rep["splitter_iterations"] <= 10
    # Errors re-measured from the reported splitters with np.searchsorted.
    runs = [np.sort(k[off[r]:off[r + 1]]) for r in range(p)]
    spl = rep["splitters"][:kw - 1]
    glb = [sum(int(np.searchsorted(r, x, side="left")) for r in runs) for x in spl]
    for i, g in enumerate(glb):
        assert abs(g - (i + 1) * n / kw) <= 0.01 / (kw - 1) * n
