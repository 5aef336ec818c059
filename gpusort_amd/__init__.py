"""Python binding of the MI355X sort-path C ABI (``include/gpusort_hip.h``).

The product is the gfx950 shared library ``libgpusort_hip.so``; this module
only loads it with ctypes and passes device pointers (torch tensors) or host
buffers (numpy arrays) through.  There is no CPU fallback: if the library is
missing, every call raises ``GpusortError``.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libgpusort_hip.so"

# gs_dtype codes (include/gpusort_hip.h)
I32, U32, F32, I64, U64, F64, NONE = 0, 1, 2, 3, 4, 5, 255
HYKSORT, SAMPLESORT, HYPERQUICKSORT = 0, 1, 2
TILES_AUTO, TILES_ARRIVAL, TILES_GROUPED = -1, 0, 1  # gs_set_tile_order
# GS_NOALLOC_SCRATCH_BYTES: workspace bytes of gs_sort_pairs_noalloc
NOALLOC_SCRATCH_BYTES = 72 << 20

STATUS_NAMES = {
    0: "GS_OK",
    1: "GS_INVALID_ARG",
    2: "GS_OUT_OF_MEMORY",
    3: "GS_ADMISSION_REFUSED",
    4: "GS_DEVICE_ERROR",
    5: "GS_NO_DEVICE",
    6: "GS_UNSUPPORTED",
    7: "GS_EMPTY_INPUT",
    8: "GS_NOT_READY",
}

_NP_DTYPE = {
    np.dtype(np.int32): I32,
    np.dtype(np.uint32): U32,
    np.dtype(np.float32): F32,
    np.dtype(np.int64): I64,
    np.dtype(np.uint64): U64,
    np.dtype(np.float64): F64,
}
_CODE_NP = {v: k for k, v in _NP_DTYPE.items()}


class GpusortError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class VrankReport(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_uint32),
        ("max_splitters", ctypes.c_uint32),
        ("n_splitters", ctypes.c_uint32),
        ("splitters", ctypes.c_void_p),
        ("bucket_sizes", ctypes.POINTER(ctypes.c_uint64)),
        ("max_bucket_sizes", ctypes.c_uint32),
        ("n_bucket_sizes", ctypes.c_uint32),
        ("final_rank_sizes", ctypes.POINTER(ctypes.c_uint64)),
        ("max_rank_error", ctypes.c_double),
        ("splitter_iterations", ctypes.c_uint32),
    ]


class KernelStat(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 48),
        ("launches", ctypes.c_uint64),
        ("total_ms", ctypes.c_double),
        ("min_ms", ctypes.c_double),
        ("max_ms", ctypes.c_double),
        ("bytes", ctypes.c_uint64),
    ]


# gs_generate distribution codes
GEN_UNIFORM_U32, GEN_UNIFORM_U64, GEN_ZIPF_U32, GEN_SORTED_U32, GEN_EQUAL_U32, GEN_NORMAL_F32, GEN_IOTA_U32, \
    GEN_IOTA_U64 = range(8)

_lib = None

# Exported C symbols and their argument types.
class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int),
        ("hbm_budget", ctypes.c_uint64),
        ("streams", ctypes.c_uint32),
        ("coalesce_max_keys", ctypes.c_uint64),
        ("batch_max_keys", ctypes.c_uint64),
        ("batch_max_requests", ctypes.c_uint32),
        ("serial", ctypes.c_int),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("n", ctypes.c_uint64),
        ("kdt", ctypes.c_int),
        ("vdt", ctypes.c_int),
        ("status", ctypes.c_int),
        ("done", ctypes.c_int),
        ("submit_us", ctypes.c_double),
        ("start_us", ctypes.c_double),
        ("done_us", ctypes.c_double),
        ("batch", ctypes.c_uint64),
        ("batch_size", ctypes.c_uint32),
        ("dev_start_us", ctypes.c_double),
        ("dev_done_us", ctypes.c_double),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "submitted", "completed", "failed", "refused", "batches", "coalesced_batches", "coalesced_requests",
        "pooled_bytes", "peak_pooled_bytes", "budget")]


class SegConfig(ctypes.Structure):
    _fields_ = [("small_max", ctypes.c_uint32), ("medium_max", ctypes.c_uint64), ("medium_geometry", ctypes.c_uint32)]


class TuneRow(ctypes.Structure):
    _fields_ = [("config", SegConfig), ("ms", ctypes.c_double)]


_SIGS = {
    "gs_get_seg_config": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SegConfig)], ctypes.c_int),
    "gs_set_seg_config": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SegConfig)], ctypes.c_int),
    "gs_autotune_segments": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
         ctypes.POINTER(SegConfig), ctypes.POINTER(TuneRow), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)],
        ctypes.c_int,
    ),
    "gs_seg_plan_create": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32, ctypes.c_int,
         ctypes.POINTER(ctypes.c_void_p)],
        ctypes.c_int,
    ),
    "gs_seg_plan_run": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "gs_seg_plan_launch_graph": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "gs_seg_plan_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "gs_multisplit": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
         ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.POINTER(ctypes.c_uint64)],
        ctypes.c_int,
    ),
    "gs_server_create": ([ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ServerConfig)], ctypes.c_int),
    "gs_server_submit": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
         ctypes.POINTER(ctypes.c_uint64)],
        ctypes.c_int,
    ),
    "gs_server_submit_host": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
         ctypes.POINTER(ctypes.c_uint64)],
        ctypes.c_int,
    ),
    "gs_server_wait": ([ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64], ctypes.c_int),
    "gs_server_drain": ([ctypes.c_void_p], ctypes.c_int),
    "gs_server_info": ([ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(RequestInfo)], ctypes.c_int),
    "gs_server_get_stats": ([ctypes.c_void_p, ctypes.POINTER(ServerStats)], ctypes.c_int),
    "gs_server_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "gs_version": ([], ctypes.c_char_p),
    "gs_last_error": ([], ctypes.c_char_p),
    "gs_get_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "gs_create": ([ctypes.POINTER(ctypes.c_void_p), ctypes.c_int], ctypes.c_int),
    "gs_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "gs_set_stream": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "gs_set_tile_order": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "gs_set_host_sort_budget": ([ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "gs_synchronize": ([ctypes.c_void_p], ctypes.c_int),
    "gs_arena_bytes": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
    "gs_release_arena": ([ctypes.c_void_p], ctypes.c_int),
    "gs_sort_keys": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int], ctypes.c_int),
    "gs_sort_pairs": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int],
        ctypes.c_int,
    ),
    "gs_sort_pairs_noalloc": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int],
        ctypes.c_int,
    ),
    "gs_sort_keys_host": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int],
        ctypes.c_int,
    ),
    "gs_sort_pairs_host": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
         ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32],
        ctypes.c_int,
    ),
    "gs_sort_segments": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_uint32, ctypes.c_int],
        ctypes.c_int,
    ),
    "gs_merge": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
         ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int],
        ctypes.c_int,
    ),
    "gs_bucket_bounds": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
         ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
        ctypes.c_int,
    ),
    "gs_vrank_sort": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(VrankReport)],
        ctypes.c_int,
    ),
    "gs_generate": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
         ctypes.c_double, ctypes.c_void_p, ctypes.c_uint32],
        ctypes.c_int,
    ),
    "gs_profile_begin": ([ctypes.c_void_p], ctypes.c_int),
    "gs_profile_end": (
        [ctypes.c_void_p, ctypes.POINTER(KernelStat), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)],
        ctypes.c_int,
    ),
    "gs_vrank_sort_host": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(VrankReport)],
        ctypes.c_int,
    ),
    "gs_gather_host": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_uint64],
        ctypes.c_int,
    ),
    "gs_rebalance_sizes": ([ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32], ctypes.c_int),
    "gs_gather": (
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
        ctypes.c_int,
    ),
}


def exported_symbols() -> list[str]:
    return sorted(_SIGS)


def load_library(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Loads ``libgpusort_hip.so`` (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise GpusortError(5, f"{p} is missing: build it with `python -m gpusort_amd.build`")
    try:  # share torch's HIP runtime when torch is present (one runtime per process)
        import torch  # noqa: F401
    except Exception:
        pass
    lib = ctypes.CDLL(str(p))
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        msg = _lib.gs_last_error().decode(errors="replace") if _lib else ""
        raise GpusortError(status, msg)


def device_count() -> int:
    lib = load_library()
    n = ctypes.c_int(0)
    _check(lib.gs_get_device_count(ctypes.byref(n)))
    return n.value


def torch_dtype_code(t) -> int:
    import torch

    m = {
        torch.int32: I32,
        getattr(torch, "uint32", None): U32,
        torch.float32: F32,
        torch.int64: I64,
        getattr(torch, "uint64", None): U64,
        torch.float64: F64,
    }
    code = m.get(t.dtype)
    if code is None:
        raise GpusortError(1, f"unsupported tensor dtype {t.dtype}")
    return code


def np_dtype_code(a: np.ndarray) -> int:
    code = _NP_DTYPE.get(a.dtype)
    if code is None:
        raise GpusortError(1, f"unsupported array dtype {a.dtype}")
    return code


def _dev_ptr(t) -> int:
    if not t.is_cuda:
        raise GpusortError(1, "expected a device tensor")
    if not t.is_contiguous():
        raise GpusortError(1, "expected a contiguous tensor")
    return t.data_ptr()


# Handles, plans and servers still open when the interpreter exits are
# closed from an atexit hook, newest first (plans and servers before the
# handle they use), while the HIP runtime is still whole; closing them later,
# from __del__ during teardown, would release device and pinned host memory
# after the runtime (or a profiler hooked into it) has shut down.
_open_objects: list = []


def _track(obj) -> None:
    if len(_open_objects) >= 64:
        _open_objects[:] = [r for r in _open_objects if r() is not None]
    _open_objects.append(weakref.ref(obj))


@atexit.register
def _close_open_objects() -> None:
    for ref in reversed(_open_objects):
        obj = ref()
        if obj is not None:
            try:
                obj.close()
            except Exception:
                pass
    _open_objects.clear()


class Handle:
    """One library handle: a device, a stream and the HBM arena."""

    def __init__(self, device: int = 0, stream=None):
        self.lib = load_library()
        h = ctypes.c_void_p()
        _check(self.lib.gs_create(ctypes.byref(h), device))
        self.h = h
        self.device = device
        _track(self)
        if stream is not None:
            self.set_stream(stream)

    def close(self):
        if self.h:
            self.lib.gs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream) -> None:
        ptr = stream if isinstance(stream, int) else getattr(stream, "cuda_stream", 0)
        _check(self.lib.gs_set_stream(self.h, ctypes.c_void_p(ptr)))

    def set_tile_order(self, order: int) -> None:
        """TILES_AUTO (grouped tiles on the null stream, arrival order
        elsewhere), TILES_ARRIVAL or TILES_GROUPED (gs_set_tile_order)."""
        _check(self.lib.gs_set_tile_order(self.h, int(order)))

    def set_host_sort_budget(self, nbytes: int) -> None:
        """Device bytes one host-memory sort may stage (gs_set_host_sort_budget);
        0 means the device's free memory.  Larger inputs are sorted in chunks
        and merged through the device."""
        _check(self.lib.gs_set_host_sort_budget(self.h, int(nbytes)))

    def use_torch_stream(self) -> None:
        import torch

        self.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    def synchronize(self) -> None:
        _check(self.lib.gs_synchronize(self.h))

    def arena_bytes(self) -> int:
        b = ctypes.c_uint64(0)
        _check(self.lib.gs_arena_bytes(self.h, ctypes.byref(b)))
        return b.value

    def release_arena(self) -> None:
        _check(self.lib.gs_release_arena(self.h))

    # ---------------------------------------------------------------- device tensors
    def sort_keys(self, keys, descending: bool = False) -> None:
        _check(self.lib.gs_sort_keys(self.h, torch_dtype_code(keys), _dev_ptr(keys), keys.numel(), int(descending)))

    def sort_pairs(self, keys, vals, descending: bool = False) -> None:
        if keys.numel() != vals.numel():
            raise GpusortError(1, "keys and values differ in length")
        _check(
            self.lib.gs_sort_pairs(
                self.h, torch_dtype_code(keys), _dev_ptr(keys), torch_dtype_code(vals), _dev_ptr(vals), keys.numel(),
                int(descending),
            )
        )

    def sort_pairs_noalloc(self, keys, vals=None, descending: bool = False) -> None:
        """gs_sort_pairs_noalloc: the sort with the library's fixed workspace
        (NOALLOC_SCRATCH_BYTES per GPU) instead of the handle's arena."""
        if vals is not None and keys.numel() != vals.numel():
            raise GpusortError(1, "keys and values differ in length")
        _check(
            self.lib.gs_sort_pairs_noalloc(
                self.h, torch_dtype_code(keys), _dev_ptr(keys), NONE if vals is None else torch_dtype_code(vals),
                None if vals is None else _dev_ptr(vals), keys.numel(), int(descending),
            )
        )

    def sort_segments(self, keys, offsets, vals=None, descending: bool = False) -> None:
        nseg = offsets.numel() - 1
        _check(
            self.lib.gs_sort_segments(
                self.h, torch_dtype_code(keys), _dev_ptr(keys), NONE if vals is None else torch_dtype_code(vals),
                None if vals is None else _dev_ptr(vals), _dev_ptr(offsets), nseg, int(descending),
            )
        )

    def seg_config(self, kdt: int, vdt: int = NONE) -> dict:
        c = SegConfig()
        _check(self.lib.gs_get_seg_config(self.h, kdt, vdt, ctypes.byref(c)))
        return {"small_max": c.small_max, "medium_max": c.medium_max, "medium_geometry": c.medium_geometry}

    def set_seg_config(self, kdt: int, vdt: int, small_max: int, medium_max: int, medium_geometry: int = 0) -> None:
        c = SegConfig(small_max, medium_max, medium_geometry)
        _check(self.lib.gs_set_seg_config(self.h, kdt, vdt, ctypes.byref(c)))

    def autotune_segments(self, kdt: int, vdt: int = NONE, total_keys: int = 1 << 25, median: float = 2048.0,
                          sigma: float = 2.0):
        """Times every size-class candidate on a synthetic batch; keeps the best in the handle."""
        best = SegConfig()
        rows = (TuneRow * 64)()
        nr = ctypes.c_uint32()
        _check(self.lib.gs_autotune_segments(self.h, kdt, vdt, total_keys, median, sigma, ctypes.byref(best), rows,
                                             64, ctypes.byref(nr)))
        table = [{"small_max": rows[i].config.small_max, "medium_max": rows[i].config.medium_max,
                  "medium_geometry": rows[i].config.medium_geometry, "ms": rows[i].ms} for i in range(nr.value)]
        return {"small_max": best.small_max, "medium_max": best.medium_max,
                "medium_geometry": best.medium_geometry}, table

    def seg_plan(self, offsets: np.ndarray, kdt: int, vdt: int = NONE, descending: bool = False) -> "SegPlan":
        return SegPlan(self, offsets, kdt, vdt, descending)

    def multisplit(self, keys, splitters: np.ndarray, out_keys, vals=None, out_vals=None, upper: bool = False,
                   descending: bool = False) -> np.ndarray:
        """Stable multisplit into len(splitters) + 1 buckets; returns the bucket offsets."""
        spl = np.ascontiguousarray(splitters, dtype=_CODE_NP[torch_dtype_code(keys)])
        offs = np.zeros(spl.size + 2, dtype=np.uint64)
        _check(self.lib.gs_multisplit(
            self.h, torch_dtype_code(keys), _dev_ptr(keys), NONE if vals is None else torch_dtype_code(vals),
            None if vals is None else _dev_ptr(vals), keys.numel(), spl.ctypes.data if spl.size else None, spl.size,
            int(upper), int(descending), _dev_ptr(out_keys), None if out_vals is None else _dev_ptr(out_vals),
            offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return offs

    def merge(self, a_keys, b_keys, out_keys, a_vals=None, b_vals=None, out_vals=None, descending=False) -> None:
        vcode = NONE if a_vals is None else torch_dtype_code(a_vals)
        _check(
            self.lib.gs_merge(
                self.h, torch_dtype_code(a_keys), _dev_ptr(a_keys), None if a_vals is None else _dev_ptr(a_vals),
                a_keys.numel(), _dev_ptr(b_keys), None if b_vals is None else _dev_ptr(b_vals), b_keys.numel(),
                vcode, _dev_ptr(out_keys), None if out_vals is None else _dev_ptr(out_vals), int(descending),
            )
        )

    def bucket_bounds(self, keys, run_off: np.ndarray, splitters: np.ndarray, upper=False, descending=False):
        run_off = np.ascontiguousarray(run_off, dtype=np.uint64)
        nruns = run_off.size - 1
        spl = np.ascontiguousarray(splitters)
        out = np.zeros(nruns * (spl.size + 2), dtype=np.uint64)
        _check(
            self.lib.gs_bucket_bounds(
                self.h, torch_dtype_code(keys), _dev_ptr(keys), run_off.ctypes.data, nruns,
                spl.ctypes.data if spl.size else None, spl.size, int(upper), int(descending), out.ctypes.data,
            )
        )
        return out.reshape(nruns, spl.size + 2)

    def vrank_sort(self, keys, rank_off: np.ndarray, vals=None, algorithm=HYKSORT, k_way=8, samples_per_rank=32,
                   descending=False, max_splitters=None, max_buckets=None):
        rank_off = np.ascontiguousarray(rank_off, dtype=np.uint64)
        p = rank_off.size - 1
        kcode = torch_dtype_code(keys)
        caps = report_capacity(p, k_way)
        if max_splitters is None:
            max_splitters = caps[0]
        if max_buckets is None:
            max_buckets = caps[1]
        spl = np.zeros(max_splitters, dtype=_CODE_NP[kcode])
        bsz = np.zeros(max_buckets, dtype=np.uint64)
        fin = np.zeros(p, dtype=np.uint64)
        rep = VrankReport()
        rep.max_splitters = max_splitters
        rep.splitters = spl.ctypes.data
        rep.bucket_sizes = bsz.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        rep.max_bucket_sizes = max_buckets
        rep.final_rank_sizes = fin.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        _check(
            self.lib.gs_vrank_sort(
                self.h, algorithm, kcode, _dev_ptr(keys), NONE if vals is None else torch_dtype_code(vals),
                None if vals is None else _dev_ptr(vals), rank_off.ctypes.data, p, k_way, samples_per_rank,
                int(descending), ctypes.byref(rep),
            )
        )
        if rep.n_splitters > max_splitters or rep.n_bucket_sizes > max_buckets:
            raise GpusortError(1, f"the report holds {rep.n_splitters} splitters and "
                               f"{rep.n_bucket_sizes} bucket sizes, more than max_splitters={max_splitters} / "
                               f"max_buckets={max_buckets} (the sort itself completed)")
        return {
            "rounds": rep.rounds,
            "splitters": spl[: rep.n_splitters].copy(),
            "bucket_sizes": bsz[: rep.n_bucket_sizes].copy(),
            "final_rank_sizes": fin.copy(),
            "max_rank_error": rep.max_rank_error,
            "splitter_iterations": rep.splitter_iterations,
        }

    def generate(self, out, dist: int, seed: int, index_base: int = 0, param: float = 0.0, cdf=None) -> None:
        """Fills a device tensor with a seeded synthetic distribution (gs_generate)."""
        cdf_arr = None if cdf is None else np.ascontiguousarray(cdf, dtype=np.float64)
        _check(
            self.lib.gs_generate(
                self.h, dist, seed & 0xFFFFFFFFFFFFFFFF, _dev_ptr(out), out.numel(), index_base, float(param),
                None if cdf_arr is None else cdf_arr.ctypes.data, 0 if cdf_arr is None else cdf_arr.size,
            )
        )

    def profile_begin(self) -> None:
        _check(self.lib.gs_profile_begin(self.h))

    def profile_end(self) -> dict:
        """Per-kernel {name: {launches, total_ms, min_ms, max_ms, bytes}} since profile_begin."""
        arr = (KernelStat * 64)()
        n = ctypes.c_uint32(0)
        _check(self.lib.gs_profile_end(self.h, arr, 64, ctypes.byref(n)))
        out = {}
        for i in range(min(n.value, 64)):
            st = arr[i]
            out[st.name.decode()] = {
                "launches": st.launches, "total_ms": st.total_ms, "min_ms": st.min_ms, "max_ms": st.max_ms,
                "bytes": st.bytes,
            }
        return out

    def gather(self, src, idx, dst) -> None:
        _check(self.lib.gs_gather(self.h, src.element_size(), _dev_ptr(src), _dev_ptr(idx), _dev_ptr(dst), dst.numel()))

    # ---------------------------------------------------------------- host arrays
    def sort_keys_host(self, keys: np.ndarray, descending: bool = False) -> None:
        if not keys.flags.c_contiguous:
            raise GpusortError(1, "expected a contiguous array")
        _check(self.lib.gs_sort_keys_host(self.h, np_dtype_code(keys), keys.ctypes.data, keys.size, int(descending)))

    def sort_pairs_host(self, keys: np.ndarray, vals: np.ndarray, descending: bool = False,
                        pairs_out: np.ndarray | None = None, pair_stride: int = 0, val_offset: int = 0) -> None:
        _check(
            self.lib.gs_sort_pairs_host(
                self.h, np_dtype_code(keys), keys.ctypes.data, np_dtype_code(vals), vals.ctypes.data, keys.size,
                int(descending), None if pairs_out is None else pairs_out.ctypes.data, pair_stride, val_offset,
            )
        )


class SegPlan:
    """Batched sort planned for fixed segment offsets (gs_seg_plan_*): run()
    queues kernel launches only; launch_graph() replays a captured HIP graph."""

    def __init__(self, handle: "Handle", offsets: np.ndarray, kdt: int, vdt: int, descending: bool):
        self.handle = handle
        self.lib = handle.lib
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        p = ctypes.c_void_p()
        _check(self.lib.gs_seg_plan_create(handle.h, kdt, vdt, off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                           off.size - 1, int(descending), ctypes.byref(p)))
        self.p = p.value
        _track(self)

    def run(self, keys, vals=None) -> None:
        _check(self.lib.gs_seg_plan_run(self.p, _dev_ptr(keys), None if vals is None else _dev_ptr(vals)))

    def launch_graph(self, keys, vals=None) -> None:
        _check(self.lib.gs_seg_plan_launch_graph(self.p, _dev_ptr(keys), None if vals is None else _dev_ptr(vals)))

    def close(self) -> None:
        if getattr(self, "p", None):
            self.lib.gs_seg_plan_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Sort server over device-resident requests (gs_server_*).  Submitted
    tensors belong to the server until their request is done."""

    def __init__(self, device: int = 0, hbm_budget: int = 200 * 10**9, streams: int = 4,
                 coalesce_max_keys: int = 1 << 16, batch_max_keys: int = 1 << 24, batch_max_requests: int = 4096,
                 serial: bool = False):
        self.lib = load_library()
        cfg = ServerConfig(device, hbm_budget, streams, coalesce_max_keys, batch_max_keys, batch_max_requests,
                           int(serial))
        p = ctypes.c_void_p()
        _check(self.lib.gs_server_create(ctypes.byref(p), ctypes.byref(cfg)))
        self.s = p.value
        _track(self)
        self._live = {}

    def submit(self, keys, vals=None, descending: bool = False) -> int:
        rid = ctypes.c_uint64()
        _check(self.lib.gs_server_submit(
            self.s, torch_dtype_code(keys), _dev_ptr(keys), NONE if vals is None else torch_dtype_code(vals),
            None if vals is None else _dev_ptr(vals), keys.numel(), int(descending), ctypes.byref(rid)))
        self._live[rid.value] = (keys, vals)  # keep the buffers alive until done
        return rid.value

    def submit_host(self, keys, vals=None, descending: bool = False) -> int:
        """Queues a request in host memory (torch CPU tensors, pinned for
        overlapping copies); sorted in place."""
        if keys.is_cuda or (vals is not None and vals.is_cuda):
            raise GpusortError(1, "submit_host takes host tensors")
        rid = ctypes.c_uint64()
        _check(self.lib.gs_server_submit_host(
            self.s, torch_dtype_code(keys), keys.data_ptr(), NONE if vals is None else torch_dtype_code(vals),
            None if vals is None else vals.data_ptr(), keys.numel(), int(descending), ctypes.byref(rid)))
        self._live[rid.value] = (keys, vals)
        return rid.value

    def wait(self, rid: int, timeout_us: int = -1) -> bool:
        st = self.lib.gs_server_wait(self.s, rid, timeout_us)
        if st == 8:
            return False
        _check(st)
        self._live.pop(rid, None)
        return True

    def drain(self) -> None:
        _check(self.lib.gs_server_drain(self.s))
        self._live.clear()

    def info(self, rid: int) -> dict:
        r = RequestInfo()
        _check(self.lib.gs_server_info(self.s, rid, ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in RequestInfo._fields_}

    def stats(self) -> dict:
        st = ServerStats()
        _check(self.lib.gs_server_get_stats(self.s, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "s", None):
            self.lib.gs_server_destroy(self.s)
            self.s = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def report_capacity(p: int, k_way: int) -> tuple[int, int]:
    """Room for a virtual-rank report: (splitters, bucket sizes).  Every
    algorithm runs at most ceil(log2 p) rounds, each with at most
    max(k, p) + 1 buckets and max(k, p) splitters per rank; capped at 2^16
    splitters and 2^20 bucket sizes.  Sized to fit rather than to the caps,
    since large zeroed host arrays cost more than the sort's tail."""
    fan = max(int(k_way), int(p), 2)
    rounds = max(1, (int(p) - 1).bit_length())
    return min(1 << 16, rounds * p * fan + p), min(1 << 20, rounds * p * (fan + 1))


def rebalance_sizes(sizes) -> np.ndarray:
    """Final rank sizes after the empty-rank rebalance (gs_rebalance_sizes)."""
    s = np.ascontiguousarray(sizes, dtype=np.uint64).copy()
    _check(load_library().gs_rebalance_sizes(s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), s.size))
    return s


def version() -> str:
    return load_library().gs_version().decode()
