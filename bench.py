#!/usr/bin/env python3
"""Benchmark of the per-rank sort path on one MI355X.

Workload (one "step"): a stable key-value sort of 2^28 uint32 keys with 2^28
uint32 payloads (payload = original index), keys uniform over [0, 2^32) from
a fixed seed, inputs resident in HBM before the timed region.  Each timed step
sorts its own pre-generated copy of the input, so every step does the same
work.  Prints ONE JSON line (rank 0, N=1).

Also reported (fewer steps, same path): keys-only 2^28 u32, the Zipf /
already-sorted / all-equal key distributions, and 2^27 float32 keys
~ N(0, 1e3); the dominant kernel's HBM roofline fraction from event timing
of every onesweep pass; and the CPU baseline (the reference's OmpUtils
merge sort on the host cores, on a bounded sample of the same input).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Gkeys/s at 2^28 u32 (% HBM roofline); batched/served p50/p99 vs 1-call"
HBM_PEAK_GBS = 8000.0
SEED = 1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def zipf_cdf(s: float = 1.1, distinct: int = 1 << 20) -> np.ndarray:
    k = np.arange(1, distinct + 1, dtype=np.float64)
    cdf = np.cumsum(np.power(k, -s))
    cdf /= cdf[-1]
    return cdf


def make_input(h, g, torch, dist: str, n: int, seed: int):
    """Device keys (and index payload) for a named distribution."""
    if dist == "normal_f32":
        keys = torch.empty(n, dtype=torch.float32, device="cuda")
        h.generate(keys, g.GEN_NORMAL_F32, seed, param=1e3)
    else:
        keys = torch.empty(n, dtype=torch.uint32, device="cuda")
        if dist == "uniform":
            h.generate(keys, g.GEN_UNIFORM_U32, seed)
        elif dist == "zipf":
            h.generate(keys, g.GEN_ZIPF_U32, seed, cdf=zipf_cdf())
        elif dist == "sorted":
            h.generate(keys, g.GEN_SORTED_U32, seed, param=float(max(1, (1 << 32) // n)))
        elif dist == "equal":
            h.generate(keys, g.GEN_EQUAL_U32, seed)
        else:
            raise ValueError(dist)
    vals = torch.empty(n, dtype=torch.uint32, device="cuda")
    h.generate(vals, g.GEN_IOTA_U32, 0)
    return keys, vals


def as_i64(torch, t):
    """Order-preserving widening for the checks (uint32 -> int64, float stays)."""
    if t.dtype == torch.uint32:
        return t.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    return t


def verify_kv(torch, k0, ks, vs) -> bool:
    """Size-independent checks: sorted keys, payload is a permutation that maps
    the input onto the output, and equal keys keep increasing payloads."""
    kk = as_i64(torch, ks)
    if not bool((kk[1:] >= kk[:-1]).all()):
        return False
    idx = as_i64(torch, vs)
    if not bool((torch.bincount(idx, minlength=idx.numel()) == 1).all()):
        return False
    if not bool((as_i64(torch, k0)[idx] == kk).all()):
        return False
    eq = kk[1:] == kk[:-1]
    return bool(((idx[1:] > idx[:-1]) | ~eq).all())


def run_workload(h, g, torch, dist: str, n: int, steps: int, warmup: int, payload: bool, profile: bool):
    keys0, vals0 = make_input(h, g, torch, dist, n, SEED)
    torch.cuda.synchronize()
    for _ in range(warmup):
        k, v = keys0.clone(), vals0.clone()
        if payload:
            h.sort_pairs(k, v)
        else:
            h.sort_keys(k)
    copies = [(keys0.clone(), vals0.clone() if payload else None) for _ in range(steps)]
    torch.cuda.synchronize()
    if profile:
        h.profile_begin()
    t0 = time.perf_counter()
    for k, v in copies:
        if payload:
            h.sort_pairs(k, v)
        else:
            h.sort_keys(k)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = h.profile_end() if profile else {}
    ok = True
    k, v = copies[-1]
    if payload:
        ok = verify_kv(torch, keys0, k, v)
    else:
        kk = as_i64(torch, k)
        ok = bool((kk[1:] >= kk[:-1]).all()) and int(kk.sum()) == int(as_i64(torch, keys0).sum())
    del copies
    torch.cuda.empty_cache()
    ms = (t1 - t0) * 1e3 / steps
    return {"gkeys_s": n / (ms * 1e-3) / 1e9, "ms_per_sort": ms, "verified": ok, "profile": prof}


def run_batched(h, g, torch, n_arrays: int, steps: int, seed: int = 4):
    """Many independent arrays: one gs_sort_segments call for the whole batch
    against one gs_sort_pairs call per array over the same device buffers.
    Lengths clip(round(lognormal(ln 2048, 2.0)), 1, 2^22), uniform u32 keys,
    u32 payload = index within the array."""
    from gpusort_amd import datagen

    lengths = datagen.lognormal_lengths(seed, n_arrays, median=2048, sigma=2.0, lo=1, hi=1 << 22)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint64)
    n = int(off[-1])
    keys0 = torch.empty(n, dtype=torch.uint32, device="cuda")
    h.generate(keys0, g.GEN_UNIFORM_U32, seed)
    starts = torch.from_numpy(off[:-1].astype(np.int64)).cuda()
    lens_t = torch.from_numpy(lengths.astype(np.int64)).cuda()
    vals0 = (torch.arange(n, dtype=torch.int64, device="cuda") - torch.repeat_interleave(starts, lens_t, output_size=n))
    vals0 = vals0.to(torch.int32).view(torch.uint32)
    del starts, lens_t
    d_off = torch.from_numpy(off.view(np.int64)).cuda()
    torch.cuda.synchronize()

    # Size classes picked by the autotuner on its own synthetic lognormal batch.
    tuned, table = h.autotune_segments(g.U32, g.U32, total_keys=1 << 26, median=2048.0, sigma=2.0)
    k, v = keys0.clone(), vals0.clone()
    h.sort_segments(k, d_off, v)  # warm-up: grows the workspace
    del k, v
    copies = [(keys0.clone(), vals0.clone()) for _ in range(steps)]
    torch.cuda.synchronize()
    h.profile_begin()
    t0 = time.perf_counter()
    for k, v in copies:
        h.sort_segments(k, d_off, v)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = h.profile_end()
    batched_ms = (t1 - t0) * 1e3 / steps
    bk, bv = copies[-1]
    del copies[:-1]

    # The same batch as a plan replayed from a captured HIP graph: kernel
    # launches only, no offset read-back, classification or map upload.
    plan = h.seg_plan(off, g.U32, g.U32)
    gk, gv = keys0.clone(), vals0.clone()
    plan.launch_graph(gk, gv)  # capture + first replay
    torch.cuda.synchronize()
    graph_same = bool(torch.equal(gk.view(torch.int32), bk.view(torch.int32))) and \
        bool(torch.equal(gv.view(torch.int32), bv.view(torch.int32)))
    graph_ms = []
    for _ in range(steps):
        gk.copy_(keys0)
        gv.copy_(vals0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.launch_graph(gk, gv)
        torch.cuda.synchronize()
        graph_ms.append((time.perf_counter() - t0) * 1e3)
    graph_same = graph_same and bool(torch.equal(gk.view(torch.int32), bk.view(torch.int32)))
    plan.close()
    del gk, gv

    # One call per array, same buffers (the calls are asynchronous: the host
    # enqueues while the device sorts).
    ok_k, ok_v = keys0.clone(), vals0.clone()
    lib = g.load_library()
    kp, vp, hh = ok_k.data_ptr(), ok_v.data_ptr(), h.h
    starts_l = [int(x) for x in off[:-1]]
    lens_l = [int(x) for x in lengths]
    fn = lib.gs_sort_pairs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, m in zip(starts_l, lens_l):
        st = fn(hh, g.U32, kp + 4 * a, g.U32, vp + 4 * a, m, 0)
        if st != 0:
            raise RuntimeError(f"gs_sort_pairs failed: {lib.gs_last_error()}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    single_ms = (t1 - t0) * 1e3

    same = bool(torch.equal(bk.view(torch.int32), ok_k.view(torch.int32))) and \
        bool(torch.equal(bv.view(torch.int32), ok_v.view(torch.int32)))
    # Spot check against numpy on the host: the largest array and 32 others.
    hk = None
    picks = [int(np.argmax(lengths))] + [int(x) for x in np.linspace(0, n_arrays - 1, 32).astype(np.int64)]
    spot = True
    for s_ in picks:
        a, b = int(off[s_]), int(off[s_ + 1])
        k_in = keys0[a:b].cpu().numpy()
        order = np.argsort(k_in, kind="stable")
        hk = bk[a:b].cpu().numpy()
        spot = spot and np.array_equal(hk, k_in[order]) and np.array_equal(bv[a:b].cpu().numpy(),
                                                                             order.astype(np.uint32))
    del copies, bk, bv, ok_k, ok_v, keys0, vals0
    torch.cuda.empty_cache()

    classes = {}
    for lo, hi, name in [(0, 1024, "<=1024"), (1025, 65536, "1025-65536"), (65537, 1 << 40, ">65536")]:
        m = (lengths >= lo) & (lengths <= hi)
        classes[name] = {"arrays": int(m.sum()), "keys_frac": round(float(lengths[m].sum()) / n, 4)}
    seg = prof.get("seg_onesweep_pass")
    pass_stats = None
    if seg and seg["launches"]:
        avg = seg["total_ms"] / seg["launches"]
        gbs = seg["bytes"] / seg["launches"] / (avg * 1e-3) / 1e9
        pass_stats = {"avg_ms": round(avg, 4), "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                      "launches": seg["launches"]}
    return {
        "arrays": n_arrays,
        "keys": n,
        "gkeys_s": round(n / (batched_ms * 1e-3) / 1e9, 3),
        "ms": round(batched_ms, 3),
        "one_call_per_array_ms": round(single_ms, 3),
        "one_call_per_array_gkeys_s": round(n / (single_ms * 1e-3) / 1e9, 3),
        "speedup_vs_one_call": round(single_ms / batched_ms, 2),
        "graph_replay_ms": round(float(np.median(graph_ms)), 3),
        "graph_replay_gkeys_s": round(n / (float(np.median(graph_ms)) * 1e-3) / 1e9, 3),
        "tuned_size_classes": tuned,
        "tuning_table_ms": [(r["small_max"], r["medium_max"], r["medium_geometry"], round(r["ms"], 3)) for r in table],
        "verified": same and spot and graph_same,
        "classes": classes,
        "seg_onesweep_pass": pass_stats,
    }


def run_served(h, g, torch, n_requests: int, seed: int = 5, loads=(0.5, 0.9, 1.5)):
    """Served sort requests through gs_server: sizes log-uniform over
    [2^10, 2^28] keys; 50% u32 keys, 30% u64 keys + u32 payload, 20% f32 keys
    ~ N(0, 1e3); device-resident inputs; Poisson arrivals.  The same requests
    go through the serial server (one call per request, one at a time: the
    one-call-per-array baseline) and the served one (4 stream slots, small
    requests coalesced).  Latency = submission -> completion seen by the
    server; every result is compared with the request sorted alone."""
    from gpusort_amd import datagen

    sizes = datagen.loguniform_sizes(seed, n_requests, lo=1 << 10, hi=1 << 28)
    u = (datagen.draws(seed + 1, n_requests) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    kinds = np.where(u < 0.5, 0, np.where(u < 0.8, 1, 2))  # 0 u32, 1 u64+u32, 2 f32
    total = int(sizes.sum())

    def fill(i, k, v):
        if kinds[i] == 0:
            h.generate(k, g.GEN_UNIFORM_U32, 1000 + i)
        elif kinds[i] == 1:
            h.generate(k, g.GEN_UNIFORM_U64, 1000 + i)
            h.generate(v, g.GEN_IOTA_U32, 0)
        else:
            h.generate(k, g.GEN_NORMAL_F32, 1000 + i, param=1e3)

    kdt = {0: torch.uint32, 1: torch.uint64, 2: torch.float32}
    bufs = []
    for i in range(n_requests):
        n = int(sizes[i])
        k = torch.empty(n, dtype=kdt[int(kinds[i])], device="cuda")
        v = torch.empty(n, dtype=torch.uint32, device="cuda") if kinds[i] == 1 else None
        bufs.append((k, v))
    expect = []
    for i, (k, v) in enumerate(bufs):
        ek = torch.empty_like(k)
        ev = None if v is None else torch.empty_like(v)
        fill(i, ek, ev)
        if ev is None:
            h.sort_keys(ek)
        else:
            h.sort_pairs(ek, ev)
        expect.append((ek, ev))
    torch.cuda.synchronize()

    def as_int(t):
        return t.view(torch.int32) if t.element_size() == 4 else t.view(torch.int64)

    def one_run(serial: bool, rate):
        for i, (k, v) in enumerate(bufs):
            fill(i, k, v)
        torch.cuda.synchronize()
        if rate is None:
            arrivals = np.zeros(n_requests)
        else:
            gaps = -np.log(1.0 - (datagen.draws(seed + 2, n_requests) >> np.uint64(11)).astype(np.float64)
                           / float(1 << 53)) / rate
            arrivals = np.cumsum(gaps) - gaps[0]
        srv = g.Server(serial=serial, streams=4, hbm_budget=200 * 10**9, coalesce_max_keys=1 << 16)
        t0 = time.perf_counter()
        ids = []
        for i, (k, v) in enumerate(bufs):
            while time.perf_counter() - t0 < arrivals[i]:
                pass
            ids.append(srv.submit(k, v))
        srv.drain()
        infos = [srv.info(r) for r in ids]
        st = srv.stats()
        srv.close()
        lat = np.array([(x["done_us"] - x["submit_us"]) * 1e-3 for x in infos])
        span_s = (max(x["done_us"] for x in infos) - min(x["submit_us"] for x in infos)) * 1e-6
        exact = sum(bool(torch.equal(as_int(k), as_int(ek))) and (v is None or bool(torch.equal(as_int(v), as_int(ev))))
                    for (k, v), (ek, ev) in zip(bufs, expect))
        return {
            "gkeys_s": round(total / span_s / 1e9, 3),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "makespan_ms": round(span_s * 1e3, 3),
            "bit_exact": f"{exact}/{n_requests}",
            "batches": int(st["batches"]),
            "coalesced_requests": int(st["coalesced_requests"]),
            "peak_pool_GB": round(st["peak_pooled_bytes"] / 1e9, 2),
        }

    out = {"requests": n_requests, "keys": total,
           "mix": {"u32": int((kinds == 0).sum()), "u64+u32": int((kinds == 1).sum()), "f32": int((kinds == 2).sum())}}
    serial_sat = one_run(True, None)
    served_sat = one_run(False, None)
    out["saturation"] = {"one_call": serial_sat, "served": served_sat,
                         "speedup": round(served_sat["gkeys_s"] / serial_sat["gkeys_s"], 2)}
    cap = n_requests / (serial_sat["makespan_ms"] * 1e-3)  # requests/s the serial server sustains
    out["loads"] = []
    for f in loads:
        rate = f * cap
        out["loads"].append({"offered_vs_one_call_capacity": f,
                             "offered_gkeys_s": round(rate * total / n_requests / 1e9, 3),
                             "one_call": one_run(True, rate), "served": one_run(False, rate)})
    del bufs, expect
    torch.cuda.empty_cache()
    return out


def run_multisplit(h, g, torch, n: int, nspl: int = 63, steps: int = 10, warmup: int = 10):
    """Stable multisplit of 2^28 uniform u32 keys + u32 payloads into 64
    buckets (splitters = evenly spaced sorted sample of the keys).  The
    warm-up calls bring the GPU to its steady clocks after the host-only
    work before this leg (the first calls after an idle spell run up to 15 %
    slower)."""
    keys, vals = make_input(h, g, torch, "uniform", n, SEED)
    sample = keys[:: max(1, n // 4096)].cpu().numpy()
    spl = np.sort(sample)[np.linspace(0, sample.size - 1, nspl + 2).astype(np.int64)[1:-1]]
    ok, ov = torch.empty_like(keys), torch.empty_like(vals)
    for _ in range(warmup):
        h.multisplit(keys, spl, ok, vals, ov)
    torch.cuda.synchronize()
    h.profile_begin()
    t0 = time.perf_counter()
    for _ in range(steps):
        offs = h.multisplit(keys, spl, ok, vals, ov)
    t1 = time.perf_counter()
    prof = h.profile_end()
    # checks: buckets hold their ranges, payloads map input to output
    idx = as_i64(torch, ov)
    kk = as_i64(torch, ok)
    ok_perm = bool((torch.bincount(idx, minlength=n) == 1).all()) and bool((as_i64(torch, keys)[idx] == kk).all())
    b = torch.from_numpy(offs.astype(np.int64)).cuda()
    seg = torch.repeat_interleave(torch.arange(nspl + 1, device="cuda"), b[1:] - b[:-1])
    spl_t = torch.from_numpy(spl.astype(np.int64)).cuda()
    lo = torch.cat([torch.tensor([-1], device="cuda"), spl_t])[seg]
    hi = torch.cat([spl_t, torch.tensor([1 << 40], device="cuda")])[seg]
    in_range = bool(((kk >= lo) & (kk < hi)).all())
    stable = bool(((idx[1:] > idx[:-1]) | (seg[1:] != seg[:-1])).all())
    st = prof.get("multisplit_pass")
    avg = st["total_ms"] / st["launches"]
    gbs = st["bytes"] / st["launches"] / (avg * 1e-3) / 1e9
    hs = prof.get("split_hist")
    del keys, vals, ok, ov, idx, kk, seg, lo, hi
    torch.cuda.empty_cache()
    return {"n": n, "buckets": nspl + 1, "ms": round((t1 - t0) * 1e3 / steps, 4),
            "pass_ms": round(avg, 4), "pass_GB/s": round(gbs, 1), "pass_frac": round(gbs / HBM_PEAK_GBS, 4),
            "pass_traffic": traffic_from_profile(n, "pmc_multisplit_kv_u32_2p28.txt"),
            "hist_ms": round(hs["total_ms"] / hs["launches"], 4) if hs else None,
            "hist_frac": round(hs["bytes"] / hs["total_ms"] * 1e3 / 1e9 / HBM_PEAK_GBS, 4) if hs else None,
            "verified": ok_perm and in_range and stable}


def run_global_index(h, g, torch, n: int, steps: int = 5):
    """Several payload vectors (the variadic HykSort overloads,
    hyksort_common_vector.inl:552-565, and ParUtils::ArrangeValues,
    par_utils.h:22-86): uniform u32 keys sorted with an int64 global index as
    the payload (stable), then three payload vectors (u32, f64, u16) gathered
    by that index.  Algorithmic bytes of one gather of E-byte elements:
    n * (8 + 2 * E) (index read, source read, destination write)."""
    keys0 = torch.empty(n, dtype=torch.uint32, device="cuda")
    h.generate(keys0, g.GEN_UNIFORM_U32, SEED + 7)
    gidx0 = torch.empty(n, dtype=torch.uint64, device="cuda")
    h.generate(gidx0, g.GEN_IOTA_U64, 0)
    srcs = [torch.arange(n, device="cuda", dtype=torch.int32).view(torch.uint32),
            torch.arange(n, device="cuda", dtype=torch.float64) * 0.5,
            (torch.arange(n, device="cuda", dtype=torch.int32) & 0xFFFF).to(torch.int16)]
    dsts = [torch.empty_like(x) for x in srcs]

    def one(k, gi):
        h.sort_pairs(k, gi)
        for src, dst in zip(srcs, dsts):
            h.gather(src, gi, dst)

    k, gi = keys0.clone(), gidx0.clone()
    one(k, gi)
    copies = [(keys0.clone(), gidx0.clone()) for _ in range(steps)]
    torch.cuda.synchronize()
    h.profile_begin()
    t0 = time.perf_counter()
    for k, gi in copies:
        one(k, gi)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = h.profile_end()
    k, gi = copies[-1]
    idx = gi.view(torch.int64)
    kk = as_i64(torch, k)
    ok = bool((kk[1:] >= kk[:-1]).all()) and bool((as_i64(torch, keys0)[idx] == kk).all())
    ok = ok and bool(((idx[1:] > idx[:-1]) | (kk[1:] != kk[:-1])).all())
    same = lambda t: t.view(torch.int32) if t.dtype == torch.uint32 else t  # noqa: E731
    ok = ok and all(bool((same(dst) == same(src)[idx]).all()) for src, dst in zip(srcs, dsts))
    gst = prof.get("gather")
    gather_ms = gst["total_ms"] / steps if gst else None
    gather_bytes = sum(n * (8 + 2 * x.element_size()) for x in srcs)
    del copies, keys0, gidx0, srcs, dsts, idx, kk
    torch.cuda.empty_cache()
    ms = (t1 - t0) * 1e3 / steps
    out = {"n": n, "payload_vectors": ["u32", "f64", "u16"], "ms": round(ms, 4),
           "gkeys_s": round(n / (ms * 1e-3) / 1e9, 3), "verified": ok}
    if gather_ms:
        gbs = gather_bytes / (gather_ms * 1e-3) / 1e9
        out.update({"gathers_ms": round(gather_ms, 4), "gather_GB/s": round(gbs, 1),
                    "gather_frac": round(gbs / HBM_PEAK_GBS, 4), "gather_algorithmic_bytes": gather_bytes,
                    "gather_traffic": traffic_from_profile(n, "pmc_gather_global_index_2p27.txt",
                                                           "traffic_bytes_three_gathers")})
    return out


def run_host_pointer(h, g, torch, n: int, reps: int = 2):
    """The host-memory form (gs_sort_pairs_host, CudaUtils::SortByKey's
    seam): copy in, sort, copy out, synchronised, on 2^28 u32 keys + u32
    payloads in pinned and in pageable host memory.  PCIe-inclusive: this
    is never the headline value."""
    keys_d = torch.empty(n, dtype=torch.uint32, device="cuda")
    h.generate(keys_d, g.GEN_UNIFORM_U32, SEED)
    src_k = keys_d.cpu()
    del keys_d
    out = {"n": n}
    for kind in ("pinned", "pageable"):
        pin = kind == "pinned"
        hk = torch.empty(n, dtype=torch.uint32, pin_memory=pin)
        hv = torch.empty(n, dtype=torch.uint32, pin_memory=pin)
        times = []
        for _ in range(reps if pin else 1):
            hk.copy_(src_k)
            hv.copy_(torch.arange(n, dtype=torch.int64).to(torch.uint32))
            t0 = time.perf_counter()
            h.sort_pairs_host(hk.numpy(), hv.numpy())
            times.append(time.perf_counter() - t0)
        k_np, v_np = hk.numpy(), hv.numpy()
        ok = bool(np.all(k_np[1:] >= k_np[:-1]))
        idx = v_np.astype(np.int64)
        ok = ok and bool(np.array_equal(src_k.numpy()[idx], k_np))
        ok = ok and bool(np.all((idx[1:] > idx[:-1]) | (k_np[1:] != k_np[:-1])))
        t = float(np.median(times))
        out[kind] = {"ms": round(t * 1e3, 2), "gkeys_s": round(n / t / 1e9, 3),
                     "GB_s_each_way": round(n * 8 / (t / 2) / 1e9, 1), "verified": ok}
        del hk, hv
    torch.cuda.empty_cache()
    return out


def run_c1(h, g, torch, with_cpu: bool):
    """Plain single-rank sort: 2^24 uniform u32 keys from seed 1, keys only;
    the GPU result must equal the reference CPU output (OmpUtils::MergeSort,
    what single-rank SampleSort / HykSort run) exactly."""
    from gpusort_amd import datagen

    n = 1 << 24
    keys = torch.empty(n, dtype=torch.uint32, device="cuda")
    h.generate(keys, g.GEN_UNIFORM_U32, SEED)
    k = keys.clone()
    h.sort_keys(k)
    copies = [keys.clone() for _ in range(10)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in copies:
        h.sort_keys(c)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / len(copies)
    out = {"n": n, "gpu_ms": round(ms, 4), "gpu_gkeys_s": round(n / (ms * 1e-3) / 1e9, 3)}
    if with_cpu:
        import oracle

        host_keys = datagen.uniform_u32(SEED, n)
        same_input = bool(np.array_equal(keys.cpu().numpy(), host_keys))
        fn = oracle.ref_merge_sort if oracle.ref_lib() is not None else oracle.merge_sort
        t0 = time.perf_counter()
        want = fn(host_keys)
        cpu_s = time.perf_counter() - t0
        out.update({"equals_reference_cpu_output": same_input and bool(np.array_equal(copies[-1].cpu().numpy(), want)),
                    "cpu_kind": "reference" if oracle.ref_lib() is not None else "port",
                    "cpu_s": round(cpu_s, 3), "cpu_cores": oracle.threads()})
    del copies, keys, k
    torch.cuda.empty_cache()
    return out


def run_c3(h, g, torch, with_cpu: bool, steps: int = 3):
    """Per-rank pipeline on one GPU: 2^27 uniform u64 keys with u32 payloads
    (original index) in 64 virtual ranks of 2^21; HykSort rounds with k = 8
    and 32 regular samples per rank.  Checked against the CPU re-run of the
    same rule (oracle): output, splitters, bucket sizes and final rank sizes;
    the CPU re-run's time is the host baseline for this configuration."""
    from gpusort_amd import datagen

    p, n = 64, 1 << 27
    off = np.arange(p + 1, dtype=np.uint64) * np.uint64(n // p)
    keys = torch.empty(n, dtype=torch.uint64, device="cuda")
    h.generate(keys, g.GEN_UNIFORM_U64, 3)
    vals = torch.empty(n, dtype=torch.uint32, device="cuda")
    h.generate(vals, g.GEN_IOTA_U32, 0)
    k, v = keys.clone(), vals.clone()
    h.vrank_sort(k, off, v, algorithm=g.HYKSORT, k_way=8, samples_per_rank=32)
    # Wall time of uninstrumented runs; the per-kernel split (HIP events
    # around every launch) from as many runs again.
    copies = [(keys.clone(), vals.clone()) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, v in copies:
        rep = h.vrank_sort(k, off, v, algorithm=g.HYKSORT, k_way=8, samples_per_rank=32)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    copies = [(keys.clone(), vals.clone()) for _ in range(steps)]
    torch.cuda.synchronize()
    h.profile_begin()
    t0 = time.perf_counter()
    for k, v in copies:
        h.vrank_sort(k, off, v, algorithm=g.HYKSORT, k_way=8, samples_per_rank=32)
    torch.cuda.synchronize()
    ms_events = (time.perf_counter() - t0) * 1e3 / steps
    prof = h.profile_end()
    out = {"p": p, "k_way": 8, "samples_per_rank": 32, "n": n, "rounds": rep["rounds"],
           "gpu_ms": round(ms, 3), "gpu_gkeys_s": round(n / (ms * 1e-3) / 1e9, 3),
           "gpu_ms_with_kernel_events": round(ms_events, 3),
           "max_rank_error": rep["max_rank_error"],
           "splitter_iterations": rep.get("splitter_iterations"),
           "kernels_ms": {name: round(st["total_ms"] / steps, 3) for name, st in prof.items()}}
    ksum = sum(st["total_ms"] for st in prof.values()) / steps
    out["kernel_sum_ms"] = round(ksum, 3)
    out["wall_over_kernel_sum"] = round(ms / ksum, 4)
    mroof = roofline_of(prof, "kway_merge_kernel")
    if mroof is not None:
        mroof["traffic"] = traffic_from_profile(n, "pmc_kway_merge_vrank_hyksort_2p27_u64_p64.txt")
    out["kway_merge_roofline"] = mroof
    if with_cpu:
        import oracle

        hk = datagen.uniform_u64(3, n)
        hv = np.arange(n, dtype=np.uint32)
        same_input = bool(np.array_equal(keys.cpu().numpy(), hk))
        t0 = time.perf_counter()
        wk, wv, wrep = oracle.vrank(hk, off.astype(np.int64), hv, algorithm="hyksort", k_way=8, samples_per_rank=32)
        cpu_s = time.perf_counter() - t0
        gk, gv = copies[-1]
        out["checks"] = {
            "input_matches_host_generator": same_input,
            "keys_and_payloads": bool(np.array_equal(gk.cpu().numpy(), wk)) and bool(np.array_equal(gv.cpu().numpy(), wv)),
            "splitters": bool(np.array_equal(rep["splitters"], wrep["splitters"])),
            "bucket_sizes": bool(np.array_equal(rep["bucket_sizes"].astype(np.int64), wrep["bucket_sizes"])),
            "final_rank_sizes": bool(np.array_equal(rep["final_rank_sizes"].astype(np.int64), wrep["final_rank_sizes"])),
        }
        out["cpu_pipeline"] = {"value": round(n / cpu_s / 1e9, 4), "unit": "Gkeys/s", "seconds": round(cpu_s, 3),
                               "cores": oracle.threads(), "kind": "port",
                               "what": "the oracle's re-run of the same per-rank rounds on the host"}
    del copies, keys, vals, k, v
    torch.cuda.empty_cache()
    return out


def run_served_host(h, g, torch, n_requests: int, seed: int = 6, loads=(0.5, 1.5)):
    """Served requests whose data is in (pinned) host memory: each request is
    copied in, sorted and copied back by the server (PCIe included).  Sizes
    log-uniform over [2^10, 2^26] keys, same dtype mix as the device-resident
    leg; serial server (one request at a time: copy in, sort, copy out) vs
    served (4 slots, copies of one slot overlapping the others' sorts, small
    requests coalesced)."""
    from gpusort_amd import datagen

    sizes = datagen.loguniform_sizes(seed, n_requests, lo=1 << 10, hi=1 << 26)
    u = (datagen.draws(seed + 1, n_requests) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    kinds = np.where(u < 0.5, 0, np.where(u < 0.8, 1, 2))
    total = int(sizes.sum())
    kdt = {0: torch.uint32, 1: torch.uint64, 2: torch.float32}
    dev_in, host, expect = [], [], []
    for i in range(n_requests):
        n = int(sizes[i])
        k = torch.empty(n, dtype=kdt[int(kinds[i])], device="cuda")
        v = None
        if kinds[i] == 0:
            h.generate(k, g.GEN_UNIFORM_U32, 2000 + i)
        elif kinds[i] == 1:
            h.generate(k, g.GEN_UNIFORM_U64, 2000 + i)
            v = torch.empty(n, dtype=torch.uint32, device="cuda")
            h.generate(v, g.GEN_IOTA_U32, 0)
        else:
            h.generate(k, g.GEN_NORMAL_F32, 2000 + i, param=1e3)
        ek, ev = k.clone(), None if v is None else v.clone()
        if ev is None:
            h.sort_keys(ek)
        else:
            h.sort_pairs(ek, ev)
        dev_in.append((k, v))
        expect.append((ek, ev))
        host.append((torch.empty(n, dtype=k.dtype, pin_memory=True),
                     None if v is None else torch.empty(n, dtype=v.dtype, pin_memory=True)))
    torch.cuda.synchronize()

    def as_int(t):
        return t.view(torch.int32) if t.element_size() == 4 else t.view(torch.int64)

    def one_run(serial: bool, rate):
        for (k, v), (hk, hv) in zip(dev_in, host):
            hk.copy_(k)
            if v is not None:
                hv.copy_(v)
        torch.cuda.synchronize()
        if rate is None:
            arrivals = np.zeros(n_requests)
        else:
            gaps = -np.log(1.0 - (datagen.draws(seed + 2, n_requests) >> np.uint64(11)).astype(np.float64)
                           / float(1 << 53)) / rate
            arrivals = np.cumsum(gaps) - gaps[0]
        srv = g.Server(serial=serial, streams=4, hbm_budget=200 * 10**9, coalesce_max_keys=1 << 16)
        t0 = time.perf_counter()
        ids = []
        for i, (hk, hv) in enumerate(host):
            while time.perf_counter() - t0 < arrivals[i]:
                pass
            ids.append(srv.submit_host(hk, hv))
        srv.drain()
        infos = [srv.info(r) for r in ids]
        srv.close()
        lat = np.array([(x["done_us"] - x["submit_us"]) * 1e-3 for x in infos])
        span_s = (max(x["done_us"] for x in infos) - min(x["submit_us"] for x in infos)) * 1e-6
        exact = 0
        for (hk, hv), (ek, ev) in zip(host, expect):
            ok = bool(torch.equal(as_int(hk), as_int(ek.cpu())))
            if hv is not None:
                ok = ok and bool(torch.equal(as_int(hv), as_int(ev.cpu())))
            exact += int(ok)
        return {"gkeys_s": round(total / span_s / 1e9, 3), "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3), "makespan_ms": round(span_s * 1e3, 3),
                "bit_exact": f"{exact}/{n_requests}"}

    out = {"requests": n_requests, "keys": total, "sizes": "log-uniform 2^10..2^26",
           "mix": {"u32": int((kinds == 0).sum()), "u64+u32": int((kinds == 1).sum()), "f32": int((kinds == 2).sum())}}
    serial_sat = one_run(True, None)
    served_sat = one_run(False, None)
    out["saturation"] = {"one_call": serial_sat, "served": served_sat,
                         "speedup": round(served_sat["gkeys_s"] / serial_sat["gkeys_s"], 2)}
    cap = n_requests / (serial_sat["makespan_ms"] * 1e-3)
    out["loads"] = []
    for f in loads:
        rate = f * cap
        out["loads"].append({"offered_vs_one_call_capacity": f,
                             "one_call": one_run(True, rate), "served": one_run(False, rate)})
    del dev_in, host, expect
    torch.cuda.empty_cache()
    return out


def roofline_of(prof: dict, name: str = "onesweep_pass"):
    st = prof.get(name)
    if not st or st["launches"] == 0:
        return None
    avg_ms = st["total_ms"] / st["launches"]
    bytes_per = st["bytes"] / st["launches"]
    achieved = bytes_per / (avg_ms * 1e-3) / 1e9
    return {
        "kernel": name,
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "bytes_per_launch": int(bytes_per),
        "avg_launch_ms": round(avg_ms, 4),
        "launches": st["launches"],
    }


def cpu_baseline(n_sample: int, reps: int = 3):
    """The reference's OmpUtils::MergeSort on (key, index) pairs (the CPU local
    sort of HykSort/SampleSort) on the host cores, over the first n_sample
    keys of the same input (all of it by default).  Checked in O(n): keys
    non-decreasing, payloads a permutation mapping input to output, ties in
    index order."""
    import oracle
    from gpusort_amd import datagen

    keys = datagen.uniform_u32(SEED, n_sample)
    vals = np.arange(n_sample, dtype=np.uint32)
    kind = "reference" if oracle.ref_lib() is not None else "port"
    fn = oracle.ref_merge_sort_pairs if kind == "reference" else oracle.merge_sort_pairs
    times = []
    out = None
    for _ in range(reps):
        out = None
        t0 = time.perf_counter()
        out = fn(keys, vals)
        times.append(time.perf_counter() - t0)
    ok_k, ok_v = out
    idx = ok_v.astype(np.int64)
    ok = bool(np.all(ok_k[1:] >= ok_k[:-1])) and bool(np.all(np.bincount(idx, minlength=n_sample) == 1))
    ok = ok and bool(np.array_equal(keys[idx], ok_k))
    ok = ok and bool(np.all((idx[1:] > idx[:-1]) | (ok_k[1:] != ok_k[:-1])))
    t = float(np.median(times))
    what = "the whole input" if n_sample == (1 << 28) else f"the first {n_sample} keys of the input"
    return {
        "value": round(n_sample / t / 1e9, 4),
        "unit": "Gkeys/s",
        "cores": oracle.threads(),
        "kind": kind,
        "sample": f"{what} (uniform u32 keys, u32 index payloads), OmpUtils::MergeSort on std::pair, "
                  f"median of {reps}",
        "seconds": round(t, 3),
        "verified": ok,
        "cpu_model": _cpu_model(),
        **_host_cores(),
        "all_usable_cores": oracle.threads() >= _host_cores()["usable_cpus"],
        "note": "cores = OpenMP threads the reference's merge sort ran with (omp_get_max_threads); "
                "usable_cpus = the CPUs this process may run on (its affinity), host_logical_cpus = all of the "
                "host's logical CPUs",
    }


def traffic_from_profile(n: int, fname: str = "pmc_onesweep_kv_u32_2p28.txt",
                         field: str = "traffic_bytes_per_launch"):
    """HBM bytes per launch of a kernel from its committed FETCH_SIZE /
    WRITE_SIZE summary under profiles/ (rocprofv3 --pmc, scaled by the
    plain-copy calibration), when it was taken on this workload size."""
    path = os.path.join(ROOT, "profiles", fname)
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^" + re.escape(field) + r"\s*=\s*(\d+)", text, re.M)
    nk = re.search(r"^n_keys\s*=\s*(\d+)", text, re.M)
    if not m or not nk or int(nk.group(1)) != n:
        return None
    return int(m.group(1))


def copy_calibration(torch, n: int, steps: int) -> None:
    """A plain device copy of n uint32 (4n bytes read, 4n written), for
    scaling the FETCH_SIZE / WRITE_SIZE counters."""
    src = torch.randint(0, 1 << 30, (n,), dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {steps} copies of {4 * n} bytes")


def _cpu_model() -> str:
    """The host CPU's full model string (lscpu's "Model name", else
    /proc/cpuinfo's "model name")."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_cores() -> dict:
    """Cores of the host: all logical CPUs, those this process may run on, and
    the OpenMP thread count the CPU baseline uses."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 0
    return {"host_logical_cpus": os.cpu_count() or 0, "usable_cpus": usable,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 28, help="keys per sort")
    ap.add_argument("--no-extra", action="store_true", help="skip the other distributions")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--arrays", type=int, default=100000, help="arrays in the batched workload")
    ap.add_argument("--requests", type=int, default=200, help="requests in the served workload")
    ap.add_argument("--cpu-n", type=int, default=1 << 28, help="CPU baseline sample size")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run plain copies of --n uint32 (counter calibration) and exit")
    ap.add_argument("--only", choices=["vrank", "batched", "multisplit", "global-index", "keys", "kv", "host"],
                    help="run one of the other configurations alone (for profiling) and print its JSON")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              f"(the per-rank path runs p virtual ranks on one device).", file=sys.stderr)
        return 2

    import torch

    import gpusort_amd as g

    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        return 1
    n = args.n
    if args.copy_calib:
        copy_calibration(torch, n, args.steps)
        return 0
    h = g.Handle(0)

    if args.only:
        log(f"[bench] only: {args.only}")
        if args.only in ("keys", "kv"):
            r = run_workload(h, g, torch, "uniform", n, args.steps, args.warmup, payload=args.only == "kv",
                             profile=True)
            roof1 = roofline_of(r["profile"])
            if roof1 is not None:
                roof1["traffic"] = traffic_from_profile(
                    n, "pmc_onesweep_kv_u32_2p28.txt" if args.only == "kv" else "pmc_onesweep_keys_u32_2p28.txt")
            r = {"gkeys_s": round(r["gkeys_s"], 3), "ms": round(r["ms_per_sort"], 4), "verified": r["verified"],
                 "roofline": roof1, "profile": r["profile"]}
        elif args.only == "vrank":
            r = run_c3(h, g, torch, not args.no_cpu)
        elif args.only == "batched":
            r = run_batched(h, g, torch, args.arrays, 3)
        elif args.only == "multisplit":
            r = run_multisplit(h, g, torch, n)
        elif args.only == "host":
            r = run_host_pointer(h, g, torch, n)
        else:
            r = run_global_index(h, g, torch, n // 2)
        print(json.dumps({"only": args.only, "result": r}), flush=True)
        h.close()
        return 0

    log(f"[bench] headline: KV u32+u32 uniform n={n} steps={args.steps} warmup={args.warmup}")
    head = run_workload(h, g, torch, "uniform", n, args.steps, args.warmup, payload=True, profile=True)
    roof = roofline_of(head["profile"])
    if roof is not None:
        roof["traffic"] = traffic_from_profile(n)
    hist = head["profile"].get("radix_hist")

    extra = {}
    if not args.no_extra:
        log("[bench] keys-only u32 uniform")
        r = run_workload(h, g, torch, "uniform", n, 3, 1, payload=False, profile=True)
        kroof = roofline_of(r["profile"])
        if kroof is not None:
            kroof["traffic"] = traffic_from_profile(n, "pmc_onesweep_keys_u32_2p28.txt")
        extra["keys_only_u32_uniform"] = {"gkeys_s": round(r["gkeys_s"], 3), "ms": round(r["ms_per_sort"], 3),
                                          "verified": r["verified"], "roofline": kroof}
        log("[bench] KV u32 uniform on a side stream (tiles handed out in arrival order)")
        side = torch.cuda.Stream()
        h.set_stream(side.cuda_stream)
        with torch.cuda.stream(side):
            r = run_workload(h, g, torch, "uniform", n, 5, 1, payload=True, profile=True)
        h.set_stream(0)
        extra["kv_u32_uniform_side_stream"] = {"gkeys_s": round(r["gkeys_s"], 3), "ms": round(r["ms_per_sort"], 3),
                                               "verified": r["verified"], "roofline": roofline_of(r["profile"]),
                                               "note": "same sort on a non-null stream: one arrival-order tile "
                                                       "ticket instead of XCD-grouped tiles"}
        for dist in ["zipf", "sorted", "equal"]:
            log(f"[bench] KV u32 {dist}")
            r = run_workload(h, g, torch, dist, n, 3, 1, payload=True, profile=False)
            extra[f"kv_u32_{dist}"] = {"gkeys_s": round(r["gkeys_s"], 3), "ms": round(r["ms_per_sort"], 3),
                                       "verified": r["verified"]}
        log("[bench] plain single-rank sort 2^24 u32 vs the reference CPU output")
        extra["single_rank_2^24_u32"] = run_c1(h, g, torch, not args.no_cpu)
        log("[bench] per-rank pipeline: 2^27 u64+u32, p=64, k=8 HykSort (+ CPU re-run)")
        extra["vrank_hyksort_2^27_u64_p64_k8"] = run_c3(h, g, torch, not args.no_cpu)
        log("[bench] global-index sort 2^27 u32 + gathers of three payload vectors")
        extra["global_index_u32_2^27_3_payloads"] = run_global_index(h, g, torch, n // 2)
        log("[bench] multisplit 2^28 u32+u32 into 64 buckets")
        extra["multisplit_kv_u32_64_buckets"] = run_multisplit(h, g, torch, n)
        log(f"[bench] batched: {args.arrays} arrays, one call vs one call per array")
        extra["batched_arrays_kv_u32"] = run_batched(h, g, torch, args.arrays, 3)
        log(f"[bench] served: {args.requests} requests, serial vs served, Poisson load sweep")
        extra["served_mixed"] = run_served(h, g, torch, args.requests)
        log("[bench] served from host memory: 100 requests, serial vs served")
        extra["served_host_memory"] = run_served_host(h, g, torch, 100)
        log("[bench] host-pointer KV sort 2^28 u32+u32 (copy in, sort, copy out)")
        extra["host_pointer_kv_u32_2^28"] = run_host_pointer(h, g, torch, n)
        log("[bench] KV f32 normal 2^27")
        r = run_workload(h, g, torch, "normal_f32", n // 2, 3, 1, payload=True, profile=False)
        extra["kv_f32_normal_2^27"] = {"gkeys_s": round(r["gkeys_s"], 3), "ms": round(r["ms_per_sort"], 3),
                                       "verified": r["verified"]}

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline n={args.cpu_n}")
        cpu = cpu_baseline(args.cpu_n)

    out = {
        "metric": METRIC,
        "value": round(head["gkeys_s"], 3),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_sort"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "plain large local sort: 2^28 uint32 keys + 2^28 uint32 payloads (original index), "
                        "stable, uniform keys from seed 1",
            "n_keys": n,
            "key": "u32",
            "payload": "u32",
            "distribution": "uniform",
        },
        "verified": head["verified"],
        "roofline": roof,
        "hist_kernel": None if not hist else {
            "avg_ms": round(hist["total_ms"] / hist["launches"], 4),
            "GB/s": round(hist["bytes"] / hist["launches"] / (hist["total_ms"] / hist["launches"] * 1e-3) / 1e9, 1),
        },
        # keys / (sum of the sort's kernel times per sort: histogram + passes), beside the wall-clock value
        "gkeys_s_kernel_sum": None if not hist or roof is None else round(
            n / ((hist["total_ms"] + roof["avg_launch_ms"] * roof["launches"]) / args.steps * 1e-3) / 1e9, 3),
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    h.close()
    return 0 if head["verified"] else 1


if __name__ == "__main__":
    sys.exit(main())
