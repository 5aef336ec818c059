"""ORACLE -- test infrastructure only.

ctypes binding of ``oracle/liboracle.so`` (the CPU restatement of the
reference's algorithms, ``oracle.cpp``) and, when built, of
``oracle/_ref/libref_omp.so`` (the reference's own OmpUtils compiled from its
source tree).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of ``bench.py`` use it, always to check results or as the
reported CPU baseline, never as the measured GPU path.

Every function follows a file:line of monotone-RK/gpusort cited in
``oracle.cpp``.  Pins (see tests/test_oracle.py): std::sort / std::stable_sort
/ std::merge / std::lower_bound known answers, the reference's own OmpUtils
(``_ref``), closed-form checkers restated from the reference's tests
(arithmetic progression, ln/log10, all-equal), and brute force over all
permutations of tiny inputs.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_omp.so"

KEY_NAMES = {
    np.dtype(np.int32): "i32",
    np.dtype(np.uint32): "u32",
    np.dtype(np.float32): "f32",
    np.dtype(np.int64): "i64",
    np.dtype(np.uint64): "u64",
    np.dtype(np.float64): "f64",
}
VAL_NAMES = {np.dtype(np.uint32): "u32", np.dtype(np.uint64): "u64", np.dtype(np.int64): "i64"}
ALGO = {"hyksort": 0, "samplesort": 1, "hyperquicksort": 2}

_lib = None
_ref = None


def build(reference: str | None = None) -> None:
    """Runs oracle/Makefile (liboracle.so, and _ref when the reference tree exists)."""
    cmd = ["make", "-s", "-C", str(HERE)]
    if reference:
        cmd.append(f"REF={reference}")
    subprocess.run(cmd, check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
    return _lib


def ref_lib() -> ctypes.CDLL | None:
    """The reference's own OmpUtils, or None when oracle/_ref was not built."""
    global _ref
    if _ref is None and REF_LIB.exists():
        _ref = ctypes.CDLL(str(REF_LIB))
    return _ref


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def set_threads(n: int) -> None:
    lib().orc_set_threads(ctypes.c_int(n))
    r = ref_lib()
    if r is not None and n > 0:
        r.ref_set_threads(ctypes.c_int(n))


def threads() -> int:
    return lib().orc_get_threads()


def merge_sort(a: np.ndarray, descending: bool = False) -> np.ndarray:
    """OmpUtils::MergeSort restated (omp_utils.h:147-207); returns a sorted copy."""
    out = np.ascontiguousarray(a).copy()
    fn = getattr(lib(), f"orc_merge_sort_{KEY_NAMES[out.dtype]}")
    fn(_p(out), ctypes.c_int64(out.size), ctypes.c_int(int(descending)))
    return out


def std_sort(a: np.ndarray, descending: bool = False) -> np.ndarray:
    out = np.ascontiguousarray(a).copy()
    fn = getattr(lib(), f"orc_std_sort_{KEY_NAMES[out.dtype]}")
    fn(_p(out), ctypes.c_int64(out.size), ctypes.c_int(int(descending)))
    return out


def ref_merge_sort(a: np.ndarray, descending: bool = False) -> np.ndarray:
    """The reference's own OmpUtils::MergeSort (oracle/_ref)."""
    r = ref_lib()
    if r is None:
        raise RuntimeError("oracle/_ref/libref_omp.so is not built")
    out = np.ascontiguousarray(a).copy()
    fn = getattr(r, f"ref_merge_sort_{KEY_NAMES[out.dtype]}")
    fn(_p(out), ctypes.c_int64(out.size), ctypes.c_int(int(descending)))
    return out


def ref_merge_sort_pairs(k: np.ndarray, v: np.ndarray, descending: bool = False):
    r = ref_lib()
    if r is None:
        raise RuntimeError("oracle/_ref/libref_omp.so is not built")
    kk = np.ascontiguousarray(k).copy()
    vv = np.ascontiguousarray(v, dtype=np.uint32).copy()
    fn = getattr(r, f"ref_merge_sort_pairs_{KEY_NAMES[kk.dtype]}_u32")
    fn(_p(kk), _p(vv), ctypes.c_int64(kk.size), ctypes.c_int(int(descending)))
    return kk, vv


def merge_sort_pairs(k: np.ndarray, v: np.ndarray, descending: bool = False):
    """CPU key-value local sort: lexicographic (key, value) (hyksort_helper.h:266-291)."""
    kk = np.ascontiguousarray(k).copy()
    vv = np.ascontiguousarray(v).copy()
    fn = getattr(lib(), f"orc_merge_sort_pairs_{KEY_NAMES[kk.dtype]}_{VAL_NAMES[vv.dtype]}")
    fn(_p(kk), _p(vv), ctypes.c_int64(kk.size), ctypes.c_int(int(descending)))
    return kk, vv


def stable_sort_pairs(k: np.ndarray, v: np.ndarray, descending: bool = False):
    """GPU key-value local sort semantics: stable by key (thrust::sort_by_key, cuda_utils.cu:477)."""
    kk = np.ascontiguousarray(k).copy()
    vv = np.ascontiguousarray(v).copy()
    fn = getattr(lib(), f"orc_stable_sort_pairs_{KEY_NAMES[kk.dtype]}_{VAL_NAMES[vv.dtype]}")
    fn(_p(kk), _p(vv), ctypes.c_int64(kk.size), ctypes.c_int(int(descending)))
    return kk, vv


def seg_sort(a: np.ndarray, offsets: np.ndarray, descending: bool = False) -> np.ndarray:
    """Every segment [off[s], off[s+1]) sorted on its own (keys only)."""
    out = np.ascontiguousarray(a).copy()
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    fn = getattr(lib(), f"orc_seg_sort_{KEY_NAMES[out.dtype]}")
    fn(_p(out), _p(off), ctypes.c_int64(off.size - 1), ctypes.c_int(int(descending)))
    return out


def seg_stable_sort_pairs(k: np.ndarray, v: np.ndarray, offsets: np.ndarray, descending: bool = False):
    """Every segment stable-sorted by key with its payload (the batched form of
    thrust::sort_by_key, cuda_utils.cu:477)."""
    kk = np.ascontiguousarray(k).copy()
    vv = np.ascontiguousarray(v).copy()
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    fn = getattr(lib(), f"orc_seg_stable_sort_pairs_{KEY_NAMES[kk.dtype]}_{VAL_NAMES[vv.dtype]}")
    fn(_p(kk), _p(vv), _p(off), ctypes.c_int64(off.size - 1), ctypes.c_int(int(descending)))
    return kk, vv


def multisplit(keys: np.ndarray, splitters: np.ndarray, upper: bool = False, descending: bool = False):
    """Stable bucket partition: (permutation, offsets of the nspl + 1 buckets)."""
    k = np.ascontiguousarray(keys)
    spl = np.ascontiguousarray(splitters, dtype=k.dtype)
    idx = np.empty(k.size, dtype=np.int64)
    offs = np.empty(spl.size + 2, dtype=np.int64)
    fn = getattr(lib(), f"orc_multisplit_{KEY_NAMES[k.dtype]}")
    fn(_p(k), ctypes.c_int64(k.size), _p(spl), ctypes.c_int64(spl.size), ctypes.c_int(int(upper)),
       ctypes.c_int(int(descending)), _p(idx), _p(offs))
    return idx, offs


def merge_lists(a: np.ndarray, b: np.ndarray, keep_low: bool, increasing: bool) -> np.ndarray:
    """MergeLists of SampleSort's small-input bitonic network (int64 lists)."""
    a = np.ascontiguousarray(a, dtype=np.int64)
    b = np.ascontiguousarray(b, dtype=np.int64)
    out = np.empty(a.size + b.size, dtype=np.int64)
    fn = lib().orc_merge_lists_i64
    fn.restype = ctypes.c_int64
    n = fn(_p(a), ctypes.c_int64(a.size), _p(b), ctypes.c_int64(b.size), ctypes.c_int(int(keep_low)),
           ctypes.c_int(int(increasing)), _p(out))
    return out[:n].copy()


def merge(a: np.ndarray, b: np.ndarray, descending: bool = False) -> np.ndarray:
    """OmpUtils::Merge restated (omp_utils.h:54-145)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.empty(a.size + b.size, dtype=a.dtype)
    fn = getattr(lib(), f"orc_merge_{KEY_NAMES[a.dtype]}")
    fn(_p(a), ctypes.c_int64(a.size), _p(b), ctypes.c_int64(b.size), _p(out), ctypes.c_int(int(descending)))
    return out


def ref_merge(a: np.ndarray, b: np.ndarray, descending: bool = False) -> np.ndarray:
    r = ref_lib()
    if r is None:
        raise RuntimeError("oracle/_ref/libref_omp.so is not built")
    a = np.ascontiguousarray(a).copy()
    b = np.ascontiguousarray(b, dtype=a.dtype).copy()
    out = np.empty(a.size + b.size, dtype=a.dtype)
    fn = getattr(r, f"ref_merge_{KEY_NAMES[a.dtype]}")
    fn(_p(a), ctypes.c_int64(a.size), _p(b), ctypes.c_int64(b.size), _p(out), ctypes.c_int(int(descending)))
    return out


def scan(a: np.ndarray, first: int = 0) -> np.ndarray:
    """OmpUtils::Scan restated (omp_utils.h:209-245): b[0]=first, b[i]=b[i-1]+a[i-1]."""
    a = np.ascontiguousarray(a, dtype=np.int64)
    b = np.zeros_like(a)
    if a.size:
        b[0] = first
    lib().orc_scan_i64(ctypes.c_int64(a.size), _p(a), _p(b))
    return b


def ref_scan(a: np.ndarray, first: int = 0) -> np.ndarray:
    r = ref_lib()
    if r is None:
        raise RuntimeError("oracle/_ref/libref_omp.so is not built")
    a = np.ascontiguousarray(a, dtype=np.int64)
    b = np.zeros_like(a)
    if a.size:
        b[0] = first
    r.ref_scan_i64(ctypes.c_int64(a.size), _p(a), _p(b))
    return b


def bounds(run: np.ndarray, splitters: np.ndarray, upper: bool = False, descending: bool = False) -> np.ndarray:
    run = np.ascontiguousarray(run)
    spl = np.ascontiguousarray(splitters, dtype=run.dtype)
    out = np.zeros(spl.size, dtype=np.int64)
    fn = getattr(lib(), f"orc_bounds_{KEY_NAMES[run.dtype]}")
    fn(_p(run), ctypes.c_int64(run.size), _p(spl), ctypes.c_int64(spl.size), ctypes.c_int(int(upper)),
       ctypes.c_int(int(descending)), _p(out))
    return out


def keep_no_empty(sizes) -> np.ndarray:
    """Final rank sizes after ParUtils::KeepNoProcessHasEmptyData (par_utils.h:103-175)."""
    s = np.ascontiguousarray(sizes, dtype=np.int64).copy()
    lib().orc_keep_no_empty(_p(s), ctypes.c_int(s.size))
    return s


def vrank(keys: np.ndarray, rank_off, vals: np.ndarray | None = None, algorithm: str = "hyksort", k_way: int = 8,
          samples_per_rank: int = 32, descending: bool = False):
    """CPU re-run of the per-rank pipeline rule; returns (keys, vals, report)."""
    kk = np.ascontiguousarray(keys).copy()
    off = np.ascontiguousarray(rank_off, dtype=np.int64)
    p = off.size - 1
    spl_cap = 1 << 16
    bcap = 1 << 20
    spl_k = np.zeros(spl_cap, dtype=kk.dtype)
    bsz = np.zeros(bcap, dtype=np.int64)
    fin = np.zeros(p, dtype=np.int64)
    n_spl = ctypes.c_int64(0)
    n_b = ctypes.c_int64(0)
    err = ctypes.c_double(0)
    rounds = ctypes.c_int(0)
    if vals is None:
        fn = getattr(lib(), f"orc_vrank_{KEY_NAMES[kk.dtype]}")
        fn(ctypes.c_int(ALGO[algorithm]), _p(kk), _p(off), ctypes.c_int(p), ctypes.c_int(k_way),
           ctypes.c_int(samples_per_rank), ctypes.c_int(int(descending)), _p(spl_k), ctypes.c_int64(spl_cap),
           ctypes.byref(n_spl), _p(bsz), ctypes.c_int64(bcap), ctypes.byref(n_b), _p(fin), ctypes.byref(err),
           ctypes.byref(rounds))
        vv = None
    else:
        vv = np.ascontiguousarray(vals).copy()
        spl_v = np.zeros(spl_cap, dtype=vv.dtype)
        fn = getattr(lib(), f"orc_vrank_pairs_{KEY_NAMES[kk.dtype]}_{VAL_NAMES[vv.dtype]}")
        fn(ctypes.c_int(ALGO[algorithm]), _p(kk), _p(vv), _p(off), ctypes.c_int(p), ctypes.c_int(k_way),
           ctypes.c_int(samples_per_rank), ctypes.c_int(int(descending)), _p(spl_k), _p(spl_v),
           ctypes.c_int64(spl_cap), ctypes.byref(n_spl), _p(bsz), ctypes.c_int64(bcap), ctypes.byref(n_b), _p(fin),
           ctypes.byref(err), ctypes.byref(rounds))
    rep = {
        "rounds": rounds.value,
        "splitters": spl_k[: n_spl.value].copy(),
        "bucket_sizes": bsz[: n_b.value].copy(),
        "final_rank_sizes": fin,
        "max_rank_error": err.value,
        "splitter_iterations": int(lib().orc_last_splitter_iterations()),
    }
    return kk, vv, rep
