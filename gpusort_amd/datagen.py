"""Seeded synthetic inputs for the sort workloads.

One counter-based generator serves every workload: the i-th 64-bit draw of
stream ``seed`` is the splitmix64 output for state ``seed + (i + 1) * G``
(G = 0x9E3779B97F4A7C15), so any slice can be generated independently and the
same arrays reach the GPU path, the CPU oracle and the CPU baseline.
Distributions:

* ``uniform_u32``  -- high 32 bits of the draw, uniform over [0, 2^32)
* ``uniform_u64``  -- the draw itself
* ``zipf_u32``     -- Zipf(s) ranks over ``distinct`` values by inverse CDF,
                      mapped through a fixed bijective 32-bit mix so that every
                      radix digit carries entropy
* ``sorted_u32``   -- ``i * step + (draw mod step)`` with step = 2^32 // n
* ``equal_u32``    -- one constant
* ``normal_f32``   -- N(0, sigma) by Box-Muller, rounded to float32, -0.0 -> +0.0
* ``lognormal_lengths`` -- segment lengths round(exp(ln(median) + sigma * z))
                      clipped to [lo, hi]
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= _M1
    z ^= z >> np.uint64(27)
    z *= _M2
    z ^= z >> np.uint64(31)
    return z


def draws(seed: int, n: int, start: int = 0, chunk: int = 1 << 24) -> np.ndarray:
    """64-bit draws start .. start+n-1 of stream ``seed``."""
    out = np.empty(n, dtype=np.uint64)
    s = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        for lo in range(0, n, chunk):
            hi = min(n, lo + chunk)
            i = np.arange(start + lo + 1, start + hi + 1, dtype=np.uint64)
            out[lo:hi] = _mix64(s + i * GOLDEN)
    return out


def uniform_u32(seed: int, n: int) -> np.ndarray:
    return (draws(seed, n) >> np.uint64(32)).astype(np.uint32)


def uniform_u64(seed: int, n: int) -> np.ndarray:
    return draws(seed, n)


def _mix32(x: np.ndarray) -> np.ndarray:
    """Bijective 32-bit mix (xorshift-multiply rounds; odd multipliers)."""
    x = x.astype(np.uint32)
    with np.errstate(over="ignore"):
        x ^= x >> np.uint32(16)
        x *= np.uint32(0x7FEB352D)
        x ^= x >> np.uint32(15)
        x *= np.uint32(0x846CA68B)
        x ^= x >> np.uint32(16)
    return x


def zipf_u32(seed: int, n: int, s: float = 1.1, distinct: int = 1 << 20) -> np.ndarray:
    k = np.arange(1, distinct + 1, dtype=np.float64)
    w = np.power(k, -s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    u = (draws(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    rank = np.searchsorted(cdf, u, side="right").astype(np.uint64) + np.uint64(1)
    rank = np.minimum(rank, np.uint64(distinct))
    return _mix32(rank.astype(np.uint32))


def sorted_u32(seed: int, n: int) -> np.ndarray:
    step = max(1, (1 << 32) // max(n, 1))
    i = np.arange(n, dtype=np.uint64)
    r = draws(seed, n) % np.uint64(step)
    return (i * np.uint64(step) + r).astype(np.uint32)


def equal_u32(seed: int, n: int) -> np.ndarray:
    c = int(draws(seed, 1)[0] >> np.uint64(32))
    return np.full(n, c, dtype=np.uint32)


def normal_f32(seed: int, n: int, sigma: float = 1e3) -> np.ndarray:
    d = draws(seed, 2 * n)
    u1 = ((d[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / (1 << 53))
    u2 = (d[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2) * sigma
    f = z.astype(np.float32)
    f[f == 0] = np.float32(0.0)  # canonical +0.0
    return f


def std_normal(seed: int, n: int) -> np.ndarray:
    d = draws(seed, 2 * n)
    u1 = ((d[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / (1 << 53))
    u2 = (d[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def lognormal_lengths(seed: int, count: int, median: float = 2048.0, sigma: float = 2.0, lo: int = 1,
                      hi: int = 1 << 22) -> np.ndarray:
    z = std_normal(seed, count)
    L = np.rint(np.exp(np.log(median) + sigma * z))
    return np.clip(L, lo, hi).astype(np.int64)


def loguniform_sizes(seed: int, count: int, lo: int = 1 << 10, hi: int = 1 << 28) -> np.ndarray:
    u = (draws(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def keys(dist: str, seed: int, n: int) -> np.ndarray:
    """Named distribution of the large-sort workload."""
    if dist == "uniform":
        return uniform_u32(seed, n)
    if dist == "zipf":
        return zipf_u32(seed, n)
    if dist == "sorted":
        return sorted_u32(seed, n)
    if dist == "equal":
        return equal_u32(seed, n)
    if dist == "normal_f32":
        return normal_f32(seed, n)
    if dist == "uniform_u64":
        return uniform_u64(seed, n)
    raise ValueError(f"unknown distribution {dist}")
