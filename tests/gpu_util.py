"""Helpers shared by the GPU tests: host <-> device copies and exact checks."""
from __future__ import annotations

import sys

import numpy as np


def dev(a: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t) -> np.ndarray:
    return t.cpu().numpy()


def say(*args) -> None:
    """Flushed progress line naming the case about to run on the GPU."""
    print(*args, file=sys.stderr, flush=True)


def assert_bits_equal(got: np.ndarray, want: np.ndarray, what: str = "") -> None:
    assert got.dtype == want.dtype, f"{what}: dtype {got.dtype} != {want.dtype}"
    assert got.shape == want.shape, f"{what}: shape {got.shape} != {want.shape}"
    g = got.view(np.uint8)
    w = want.view(np.uint8)
    if not np.array_equal(g, w):
        gi = got.view(np.uint32 if got.itemsize == 4 else np.uint64)
        wi = want.view(np.uint32 if want.itemsize == 4 else np.uint64)
        bad = np.nonzero(gi != wi)[0]
        raise AssertionError(f"{what}: {bad.size} of {got.size} elements differ, first at {bad[:5]}: "
                             f"got {got[bad[:5]]} want {want[bad[:5]]}")


def stable_order(keys: np.ndarray, descending: bool = False) -> np.ndarray:
    """Permutation of a stable sort by key (numpy's stable sort, independent of the oracle)."""
    if not descending:
        return np.argsort(keys, kind="stable")
    if keys.dtype.kind == "u":
        inv = ~keys
    elif keys.dtype.kind == "i":
        inv = ~keys  # bitwise not reverses signed order too
    else:
        inv = -keys
    return np.argsort(inv, kind="stable")
