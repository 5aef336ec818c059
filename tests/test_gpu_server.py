"""Sort server (gs_server_*): every request's result equals the same input
sorted alone by gs_sort_keys / gs_sort_pairs, whether it ran coalesced with
other small requests, concurrently with large ones, or in the serial
one-call-per-request mode; admission respects the HBM budget."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import gpusort_amd as g
from gpusort_amd import datagen
from gpu_util import say

pytestmark = pytest.mark.gpu

KINDS = [("u32", None), ("u64", torch.uint32), ("f32", None), ("u32", torch.uint32)]


def make_request(h, kind, n, seed):
    kname, vdt = kind
    if kname == "u32":
        k = torch.empty(n, dtype=torch.uint32, device="cuda")
        h.generate(k, g.GEN_UNIFORM_U32, seed)
    elif kname == "u64":
        k = torch.empty(n, dtype=torch.uint64, device="cuda")
        h.generate(k, g.GEN_UNIFORM_U64, seed)
    else:
        k = torch.empty(n, dtype=torch.float32, device="cuda")
        h.generate(k, g.GEN_NORMAL_F32, seed, param=1e3)
    v = None
    if vdt is not None:
        v = torch.empty(n, dtype=vdt, device="cuda")
        h.generate(v, g.GEN_IOTA_U32, 0)
    return k, v


def sizes_mix(seed, count):
    # log-uniform 2^4 .. 2^22, plus edge sizes
    s = datagen.loguniform_sizes(seed, count, lo=1 << 4, hi=1 << 22).tolist()
    return s + [0, 1, 2, 4096, 4097, 16384, 16385, 65536, 65537]


def same(a, b) -> bool:
    if a is None:
        return b is None
    ia = a.view(torch.int32) if a.element_size() == 4 else a.view(torch.int64)
    ib = b.view(torch.int32) if b.element_size() == 4 else b.view(torch.int64)
    return bool(torch.equal(ia, ib))


def run_mix(handle, server, seed, descending_every=5):
    sizes = sizes_mix(seed, 60)
    reqs = []
    for i, n in enumerate(sizes):
        kind = KINDS[i % len(KINDS)]
        desc = (i % descending_every) == 0
        k, v = make_request(handle, kind, int(n), seed * 1000 + i)
        reqs.append((kind, int(n), desc, k, v))
    torch.cuda.synchronize()
    say(f"server mix seed={seed} requests={len(reqs)} keys={sum(r[1] for r in reqs)}")
    ids = [server.submit(k, v, desc) for (_, _, desc, k, v) in reqs]
    server.drain()
    for rid, (kind, n, desc, k, v) in zip(ids, reqs):
        info = server.info(rid)
        assert info["done"] == 1 and info["status"] == 0
        assert info["submit_us"] <= info["done_us"]
        k0, v0 = make_request(handle, kind, n, seed * 1000 + ids.index(rid))
        if v0 is None:
            handle.sort_keys(k0, descending=desc)
        else:
            handle.sort_pairs(k0, v0, descending=desc)
        torch.cuda.synchronize()
        assert same(k, k0), f"request {rid} ({kind}, n={n}, desc={desc}) keys differ from the alone sort"
        assert same(v, v0), f"request {rid} payloads differ"
    return server.stats()


def test_served_mix_equals_alone(handle):
    with g.Server(streams=4, coalesce_max_keys=1 << 16) as srv:
        st = run_mix(handle, srv, 1)
    assert st["completed"] == st["submitted"] and st["failed"] == 0
    assert st["coalesced_batches"] > 0 and st["batches"] < st["submitted"]
    assert st["peak_pooled_bytes"] <= st["budget"]


def test_serial_mode_equals_alone(handle):
    with g.Server(serial=True) as srv:
        st = run_mix(handle, srv, 2)
    assert st["coalesced_batches"] == 0
    assert st["batches"] == st["submitted"] - 2  # sizes 0 and 1 complete at submission


def test_tight_budget_still_completes(handle):
    # Room for the workspace of one 2^22-key KV sort at a time.
    budget = 80 << 20
    with g.Server(streams=4, hbm_budget=budget) as srv:
        st = run_mix(handle, srv, 3)
        assert st["peak_pooled_bytes"] <= budget
        big = torch.empty(1 << 26, dtype=torch.uint32, device="cuda")
        with pytest.raises(g.GpusortError) as e:
            srv.submit(big, torch.empty_like(big))
        assert e.value.status == 3  # GS_ADMISSION_REFUSED
        assert srv.stats()["refused"] == 1


def test_wait_and_timeout(handle):
    with g.Server(streams=2) as srv:
        k, v = make_request(handle, ("u32", torch.uint32), 1 << 20, 77)
        rid = srv.submit(k, v)
        assert srv.wait(rid, timeout_us=-1)
        info = srv.info(rid)
        assert info["submit_us"] <= info["start_us"] <= info["done_us"]
        k0, v0 = make_request(handle, ("u32", torch.uint32), 1 << 20, 77)
        handle.sort_pairs(k0, v0)
        torch.cuda.synchronize()
        assert same(k, k0) and same(v, v0)


def pinned_copy(t):
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t)
    return h


@pytest.mark.parametrize("serial", [False, True], ids=["served", "serial"])
def test_host_requests_equal_alone(handle, serial):
    sizes = sizes_mix(4, 40)
    reqs = []
    for i, n in enumerate(sizes):
        kind = KINDS[i % len(KINDS)]
        k, v = make_request(handle, kind, int(n), 4000 + i)
        torch.cuda.synchronize()
        hk = pinned_copy(k)
        hv = None if v is None else pinned_copy(v)
        desc = i % 3 == 0
        if v is None:
            handle.sort_keys(k, descending=desc)
        else:
            handle.sort_pairs(k, v, descending=desc)
        reqs.append((hk, hv, k, v, desc))
    torch.cuda.synchronize()
    say(f"host requests serial={serial} count={len(reqs)}")
    with g.Server(streams=4, serial=serial) as srv:
        ids = [srv.submit_host(hk, hv, desc) for (hk, hv, _, _, desc) in reqs]
        srv.drain()
        st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == len(reqs)
    for rid, (hk, hv, k, v, desc) in zip(ids, reqs):
        assert same(hk, k.cpu()), f"host request {rid} keys differ from the alone sort"
        assert same(hv, None if v is None else v.cpu()), f"host request {rid} payloads differ"


def test_host_and_device_requests_mixed(handle):
    reqs = []
    for i, n in enumerate(sizes_mix(5, 30)):
        kind = KINDS[i % len(KINDS)]
        k, v = make_request(handle, kind, int(n), 5000 + i)
        k0, v0 = k.clone(), None if v is None else v.clone()
        if v is None:
            handle.sort_keys(k0)
        else:
            handle.sort_pairs(k0, v0)
        torch.cuda.synchronize()
        on_host = i % 2 == 1
        if on_host:
            k, v = pinned_copy(k), None if v is None else pinned_copy(v)
        reqs.append((on_host, k, v, k0, v0))
    say(f"mixed host/device requests count={len(reqs)}")
    with g.Server(streams=3) as srv:
        for on_host, k, v, _, _ in reqs:
            (srv.submit_host if on_host else srv.submit)(k, v)
        srv.drain()
    for on_host, k, v, k0, v0 in reqs:
        assert same(k if on_host else k.cpu(), k0.cpu())
        if v is not None:
            assert same(v if on_host else v.cpu(), v0.cpu())


def test_slots_overlap_on_the_device(handle):
    # Two large requests submitted together go to two slots; each slot has
    # its own stream, so on the device clock the second batch starts before
    # the first completes (on one shared stream it would start after).
    n = 1 << 26
    with g.Server(streams=2, coalesce_max_keys=1 << 10) as srv:
        a = make_request(handle, ("u32", torch.uint32), n, 81)
        b = make_request(handle, ("u64", None), n, 82)
        torch.cuda.synchronize()
        say(f"slot overlap: two requests of {n} keys")
        ra = srv.submit(*a)
        rb = srv.submit(*b)
        srv.drain()
        ia, ib = srv.info(ra), srv.info(rb)
    assert ia["batch"] != ib["batch"]
    for i in (ia, ib):
        assert i["status"] == 0 and 0 <= i["dev_start_us"] < i["dev_done_us"], i
    assert max(ia["dev_start_us"], ib["dev_start_us"]) < min(ia["dev_done_us"], ib["dev_done_us"]), (ia, ib)
    for kind, req, seed in ((("u32", torch.uint32), a, 81), (("u64", None), b, 82)):
        k0, v0 = make_request(handle, kind, n, seed)
        if v0 is None:
            handle.sort_keys(k0)
        else:
            handle.sort_pairs(k0, v0)
        torch.cuda.synchronize()
        assert same(req[0], k0) and same(req[1], v0)
