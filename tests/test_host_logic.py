"""Host-side logic of the package that needs no device: the room the
virtual-rank wrapper gives a report must hold every report the rounds can
produce (the oracle's re-run of the same rule gives the actual counts)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from gpusort_amd import datagen, report_capacity


@pytest.mark.parametrize("algorithm", ["hyksort", "samplesort", "hyperquicksort"])
def test_report_capacity_holds_every_report(algorithm):
    for p in [1, 2, 3, 4, 5, 6, 8, 12, 16, 31, 32, 48, 64, 128]:
        for k in ([2, 4, 8, 16, 64] if algorithm == "hyksort" else [8]):
            per_rank = 5 * p + 7  # past the small-input network (N >= 5 p^2)
            n = per_rank * p
            off = np.arange(p + 1, dtype=np.int64) * per_rank
            keys = (datagen.draws(p * 131 + k, n) >> np.uint64(40)).astype(np.uint32)
            vals = np.arange(n, dtype=np.uint32)
            _, _, rep = oracle.vrank(keys, off, vals, algorithm=algorithm, k_way=k, samples_per_rank=8)
            spl_cap, bkt_cap = report_capacity(p, k)
            assert len(rep["splitters"]) <= spl_cap, (algorithm, p, k)
            assert len(rep["bucket_sizes"]) <= bkt_cap, (algorithm, p, k)


def test_report_capacity_caps():
    assert report_capacity(1, 8) == (9, 9)
    spl, bkt = report_capacity(64, 8)
    assert spl == 6 * 64 * 64 + 64 and bkt == 6 * 64 * 65
    assert report_capacity(4096, 8) == (1 << 16, 1 << 20)
