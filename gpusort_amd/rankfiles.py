"""Per-rank data files in the reference's on-disk format, and a harness that
sorts a directory of them on the GPU.

Format (src/test/utils.h:610-661, src/test/ifstream.h, src/test/ofstream.h):
``<dir>/<rank><ext>`` = int64 element count + the raw elements.  Keys-only
data uses ``.dat``; key-value data stores keys in ``.key`` and value vector i
(1-based) in ``.v<i>``.  ``save_data`` cuts an array into ``np`` blocks of
``calc_num_elems_each_block(total, np)`` elements (include/common.h:146-154)
and, like the reference, leaves the tail unwritten when that rule falls short.

Harness (the reference's test drivers load rank files, sort and validate,
src/test/utils.h:815-920)::

    python -m gpusort_amd.rankfiles --data DIR --np P --type uint32 \\
        [--values float64 int32 ...] [--algo hyksort --k 8] [--descending] --out OUTDIR

reads the ranks, sorts them with the per-rank pipeline on the GPU (one value
vector: lexicographic (key, value); several: keys with their global index,
then every value vector gathered by index), writes each rank's result in the
same format and checks the keys with the reference's CheckValid
(src/test/utils.h:324-387).  Exit status 0 when valid.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

KEY_EXT = ".key"
VALUE_EXT_PREFIX = ".v"

TYPES = {
    "int32": np.int32, "int": np.int32, "uint32": np.uint32, "unsigned": np.uint32, "float32": np.float32,
    "float": np.float32, "int64": np.int64, "uint64": np.uint64, "float64": np.float64, "double": np.float64,
}


def calc_num_elems_each_block(total: int, p: int) -> int:
    r = total % p
    if r == 0:
        return total // p
    tmp = (total + (p - r)) // p
    return tmp - 1 if tmp * (p - 1) >= total else tmp


def file_path(parent: str, rank: int, ext: str = ".dat") -> str:
    return os.path.join(parent, f"{rank}{ext}")


def read_data(rank: int, parent: str, ext: str, dtype) -> np.ndarray:
    path = file_path(parent, rank, ext)
    with open(path, "rb") as f:
        head = f.read(8)
        if len(head) != 8:
            raise ValueError(f"truncated data file {path}")
        n = int(np.frombuffer(head, dtype=np.int64)[0])
        data = np.fromfile(f, dtype=dtype, count=n)
    if data.size != n:
        raise ValueError(f"truncated data file {path}")
    return data


def save_my_data(rank: int, parent: str, ext: str, data: np.ndarray) -> None:
    with open(file_path(parent, rank, ext), "wb") as f:
        f.write(np.int64(data.size).tobytes())
        np.ascontiguousarray(data).tofile(f)


def save_data(np_: int, parent: str, ext: str, data: np.ndarray) -> list[int]:
    """Writes the reference's block partition; returns the sizes written."""
    total = data.size
    each = calc_num_elems_each_block(total, np_)
    sizes = []
    for i in range(np_):
        start = each * i
        end = min(start + each - 1, total - 1)
        size = max(end - start + 1, 0)
        save_my_data(i, parent, ext, data[start:start + size])
        sizes.append(size)
    return sizes


def read_ranks(np_: int, parent: str, ext: str, dtype) -> tuple[np.ndarray, np.ndarray]:
    parts = [read_data(r, parent, ext, dtype) for r in range(np_)]
    off = np.concatenate([[0], np.cumsum([p.size for p in parts])]).astype(np.uint64)
    return (np.concatenate(parts) if parts else np.zeros(0, dtype=dtype)), off


def save_ranks(off: np.ndarray, parent: str, ext: str, data: np.ndarray) -> None:
    for r in range(off.size - 1):
        save_my_data(r, parent, ext, data[int(off[r]):int(off[r + 1])])


def check_valid(parts: list[np.ndarray], increasing: bool) -> bool:
    """CheckValid across ranks: every rank ordered and the rank boundaries in
    order (src/test/utils.h:324-387); an empty rank fails."""
    for p in parts:
        if p.size == 0:
            return False
        bad = (p[:-1] > p[1:]) if increasing else (p[:-1] < p[1:])
        if bool(np.any(bad)):
            return False
    for a, b in zip(parts[:-1], parts[1:]):
        if (a[-1] > b[0]) if increasing else (a[-1] < b[0]):
            return False
    return True


def sort_rank_files(data: str, np_: int, key_type: str, value_types: list[str], algo: str, k_way: int,
                    samples: int, descending: bool, out: str) -> bool:
    import torch

    import gpusort_amd as g

    kdt = np.dtype(TYPES[key_type])
    ext = KEY_EXT if value_types else ".dat"
    keys, off = read_ranks(np_, data, ext, kdt)
    values = [read_ranks(np_, data, f"{VALUE_EXT_PREFIX}{i + 1}", np.dtype(TYPES[t]))[0]
              for i, t in enumerate(value_types)]
    for v in values:
        if v.size != keys.size:
            raise ValueError("Data corrupted.")
    algos = {"hyksort": g.HYKSORT, "samplesort": g.SAMPLESORT, "hyperquicksort": g.HYPERQUICKSORT}
    h = g.Handle(0)
    tk = torch.from_numpy(keys).cuda()
    direct = len(values) == 1 and values[0].dtype.itemsize in (4, 8)
    if direct:
        tv = torch.from_numpy(values[0]).cuda()
    elif values:
        tv = torch.arange(keys.size, dtype=torch.int64, device="cuda")
    else:
        tv = None
    rep = h.vrank_sort(tk, off, tv, algorithm=algos[algo], k_way=k_way, samples_per_rank=samples,
                       descending=descending)
    sizes = rep["final_rank_sizes"].astype(np.int64)
    new_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    sk = tk.cpu().numpy()
    if direct:
        out_values = [tv.cpu().numpy()]
    elif values:
        # ParUtils::ArrangeValues on the device: value[i] = old_value[index[i]]
        out_values = []
        for v in values:
            raw = v.view({1: np.uint8, 2: np.int16, 4: np.int32, 8: np.int64}[v.itemsize])
            src = torch.from_numpy(raw).cuda()
            dst = torch.empty_like(src)
            h.gather(src, tv, dst)
            out_values.append(dst.cpu().numpy().view(v.dtype))
    else:
        out_values = []
    os.makedirs(out, exist_ok=True)
    save_ranks(new_off, out, ext, sk)
    for i, v in enumerate(out_values):
        save_ranks(new_off, out, f"{VALUE_EXT_PREFIX}{i + 1}", v)
    h.close()
    parts = [sk[int(new_off[r]):int(new_off[r + 1])] for r in range(np_)]
    return check_valid(parts, not descending)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Sort a directory of per-rank data files on the GPU.")
    ap.add_argument("--data", required=True)
    ap.add_argument("--np", type=int, required=True)
    ap.add_argument("--type", default="uint32", choices=sorted(TYPES))
    ap.add_argument("--values", nargs="*", default=[], choices=sorted(TYPES))
    ap.add_argument("--algo", default="hyksort", choices=["hyksort", "samplesort", "hyperquicksort"])
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--samples", type=int, default=32)
    ap.add_argument("--descending", action="store_true")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    ok = sort_rank_files(a.data, a.np, a.type, a.values, a.algo, a.k, a.samples, a.descending, a.out)
    print("valid" if ok else "INVALID")
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
