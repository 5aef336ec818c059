"""Writes tests/golden/omp_utils_golden.npz from the reference's own CPU engine.

The vectors are inputs and outputs of OmpUtils::MergeSort / Merge / Scan
(include/omp/omp_utils.h:54-245 of monotone-RK/gpusort) as compiled from the
reference tree by oracle/Makefile into oracle/_ref/libref_omp.so, including
the std::pair key-value form the HykSort/SampleSort CPU branch sorts
(include/helpers/hyksort_helper.h:266-271).  Run here, where the reference
tree exists: ``python tests/golden/make_golden.py``.  The fixture lets the
oracle be checked against the reference where the tree is absent.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from gpusort_amd import datagen  # noqa: E402

OUT = Path(__file__).resolve().parent / "omp_utils_golden.npz"


def small_keys(dtype, n, seed, distinct):
    r = datagen.draws(seed, n) % np.uint64(distinct)
    d = np.dtype(dtype)
    if d.kind == "f":
        return (r.astype(np.float64) * 0.25 - distinct / 8).astype(d)
    if d.kind == "i":
        return (r.astype(np.int64) - distinct // 2).astype(d)
    return r.astype(d)


def main() -> int:
    oracle.build(reference="/root/reference")
    if oracle.ref_lib() is None:
        print("reference tree not available: oracle/_ref not built", file=sys.stderr)
        return 1
    out = {}
    cases = 0
    for threads in (1, 3, 8):
        oracle.set_threads(threads)
        for dtype in (np.uint32, np.int32, np.float32, np.uint64, np.int64, np.float64):
            for n, distinct in ((0, 1), (1, 1), (17, 5), (600, 50), (1201, 1 << 30)):
                for desc in (0, 1):
                    k = small_keys(dtype, n, 1000 * threads + n, distinct)
                    tag = f"sort_t{threads}_{np.dtype(dtype).name}_n{n}_d{desc}"
                    out[tag + "_in"] = k
                    out[tag + "_out"] = oracle.ref_merge_sort(k, bool(desc))
                    cases += 1
        for n in (0, 5, 999, 2500):
            k = small_keys(np.uint32, n, 7 + n, 40)
            v = datagen.draws(99 + n, n).astype(np.uint32) % np.uint32(1000)
            for desc in (0, 1):
                tag = f"pairs_t{threads}_n{n}_d{desc}"
                rk, rv = oracle.ref_merge_sort_pairs(k, v, bool(desc))
                out[tag + "_kin"], out[tag + "_vin"], out[tag + "_kout"], out[tag + "_vout"] = k, v, rk, rv
                cases += 1
        for na, nb in ((0, 0), (0, 9), (9, 0), (300, 1234), (1024, 1024)):
            a = np.sort(small_keys(np.int64, na, na + 1, 60))
            b = np.sort(small_keys(np.int64, nb, nb + 2, 60))
            tag = f"merge_t{threads}_{na}_{nb}"
            out[tag + "_a"], out[tag + "_b"], out[tag + "_out"] = a, b, oracle.ref_merge(a, b)
            cases += 1
        for cnt in (1, 10, 99, 2000):
            a = (datagen.draws(cnt, cnt) % np.uint64(100)).astype(np.int64)
            tag = f"scan_t{threads}_{cnt}"
            out[tag + "_a"], out[tag + "_out"] = a, oracle.ref_scan(a, first=3)
            cases += 1
    np.savez_compressed(OUT, **out)
    print(f"wrote {cases} cases to {OUT} ({OUT.stat().st_size} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
