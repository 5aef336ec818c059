"""Host-memory sorts of inputs larger than the device budget: chunks sorted on
the device, then a binary tree of stable 2-way merges streamed through the
device in blocks (the reference's CudaUtils::Sort / SortByKey with
SplitChunks and ParallelMergeArrays, src/cuda_utils.cu:97-243, :323-371,
:508-520).  A small budget set on the handle forces the chunked path at sizes
the oracle checks in seconds; every result is compared bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import gpusort_amd as g
import oracle
from gpusort_amd import datagen
from tests.gpu_util import assert_bits_equal, say
from tests.test_gpu_sort import make_keys

pytestmark = pytest.mark.gpu

KEY_TYPES = [np.uint32, np.int32, np.float32, np.uint64, np.int64, np.float64]
CHUNK = 20_000


@pytest.fixture
def budget(handle):
    """Sets the handle's budget to `chunk` elements of the given widths;
    restores the default (free device memory) afterwards."""

    def set_chunk(chunk: int, kb: int, vb: int = 0) -> None:
        handle.set_host_sort_budget(chunk * 3 * (kb + vb))

    yield set_chunk
    handle.set_host_sort_budget(0)


def _name(t):
    return np.dtype(t).name


@pytest.mark.parametrize("desc", [False, True], ids=["asc", "desc"])
@pytest.mark.parametrize("kdt", KEY_TYPES, ids=_name)
def test_chunked_keys(handle, budget, kdt, desc):
    """6 chunks (3 merge levels, one run without a partner at level 2)."""
    n = 5 * CHUNK + 1234
    keys = np.ascontiguousarray(make_keys(kdt, n, seed=n + int(desc)))
    budget(CHUNK, np.dtype(kdt).itemsize)
    say(f"chunked keys {_name(kdt)} desc={desc} n={n}")
    k = keys.copy()
    handle.sort_keys_host(k, descending=desc)
    assert_bits_equal(k, oracle.merge_sort(keys, descending=desc), "keys")


# (key type, value type, descending, chunks, last chunk's length, distinct keys or None)
PAIR_CASES = [
    (np.uint32, np.uint32, False, 2, 777, None),
    (np.int64, np.uint32, True, 4, CHUNK, 50),  # 2 levels: chunks sorted in place
    (np.float32, np.uint64, False, 5, 1, None),
    (np.uint64, np.uint64, True, 9, 4321, 1000),  # 4 levels
    (np.int32, np.uint32, False, 3, 10, 1),  # every key equal: pure stability
    (np.float64, np.uint32, True, 6, CHUNK - 1, 7),
]


@pytest.mark.parametrize("case", PAIR_CASES,
                         ids=lambda c: f"{_name(c[0])}-{_name(c[1])}-{'desc' if c[2] else 'asc'}-c{c[3]}"
                         f"{'-d%d' % c[5] if c[5] else ''}")
def test_chunked_pairs_stable(handle, budget, case):
    """Key-value sorts are stable by key across chunks: ties keep the input
    order, as CudaUtils::SortByKey's thrust::sort_by_key does."""
    kdt, vdt, desc, chunks, last, distinct = case
    n = (chunks - 1) * CHUNK + last
    keys = make_keys(kdt, n, seed=chunks * 31 + last)
    if distinct:
        keys = keys[datagen.draws(n + 5, n) % np.uint64(distinct)]
    keys = np.ascontiguousarray(keys)
    vals = np.arange(n, dtype=np.uint64).astype(vdt)
    budget(CHUNK, np.dtype(kdt).itemsize, np.dtype(vdt).itemsize)
    say(f"chunked pairs {case} n={n}")
    k, v = keys.copy(), vals.copy()
    handle.sort_pairs_host(k, v, descending=desc)
    wk, wv = oracle.stable_sort_pairs(keys, vals, descending=desc)
    assert_bits_equal(k, wk, "keys")
    assert_bits_equal(v, wv, "payloads")


def test_chunked_pairs_out(handle, budget):
    """The pair records (std::pair<int64_t, uint32_t> layout) of a chunked
    sort equal the sorted keys and values."""
    n = 3 * CHUNK + 99
    keys = datagen.uniform_u64(12, n).view(np.int64)
    vals = np.arange(n, dtype=np.uint32)
    pairs = np.zeros(n, dtype=np.dtype([("k", np.int64), ("v", np.uint32), ("pad", np.uint32)]))
    budget(CHUNK, 8, 4)
    k, v = keys.copy(), vals.copy()
    handle.sort_pairs_host(k, v, pairs_out=pairs, pair_stride=16, val_offset=8)
    wk, wv = oracle.stable_sort_pairs(keys, vals)
    assert_bits_equal(k, wk, "keys")
    assert_bits_equal(v, wv, "payloads")
    assert np.array_equal(pairs["k"], wk) and np.array_equal(pairs["v"], wv)


def test_chunked_matches_one_piece(handle, budget):
    """The same input sorted in one piece (default budget) and in 7 chunks."""
    n = 6 * CHUNK + 5
    keys = datagen.uniform_u32(13, n)
    vals = np.arange(n, dtype=np.uint32)
    k1, v1 = keys.copy(), vals.copy()
    handle.sort_pairs_host(k1, v1)
    budget(CHUNK, 4, 4)
    k2, v2 = keys.copy(), vals.copy()
    handle.sort_pairs_host(k2, v2)
    assert np.array_equal(k1, k2) and np.array_equal(v1, v2)


def test_budget_below_one_element(handle, budget):
    handle.set_host_sort_budget(1)
    k = datagen.uniform_u32(14, 100)
    with pytest.raises(g.GpusortError) as e:
        handle.sort_keys_host(k)
    assert e.value.status == 2  # GS_OUT_OF_MEMORY, the reference's "Couldn't allocate device vector"
