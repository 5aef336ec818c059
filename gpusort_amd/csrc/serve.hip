// Sort server (gs_server_*): continuous batching of sort requests on one GPU.
//
// The reference has no serving layer: every caller runs its own sort call
// (CudaUtils::Sort / SortByKey, src/cuda_utils.cu:97-243), allocating and
// freeing device memory per call.  The server keeps N stream slots, each with
// its own stream and pooled HBM workspace (a gs_handle), and one scheduler
// thread that
//   - admits pending requests in arrival order whenever a slot is free and the
//     workspace the batch needs fits in the HBM budget (idle slots give their
//     pooled memory back when the budget is short);
//   - coalesces small requests of one (key type, payload type, direction) into
//     one segmented launch: a gather kernel copies them into a staging buffer,
//     the segmented sort runs, a scatter kernel copies them back;
//   - runs every other request as its own sort on its slot's stream (one
//     non-blocking stream per slot), so requests that arrive while others
//     run overlap on the device;
//   - marks requests done when their slot's event completes, and reads the
//     device clock of each batch's start and completion (events on the
//     slot's stream, against one reference event taken at creation).
// Requests are device-resident (sorted in place) or in host memory (copied
// into the slot's staging buffers, sorted, copied back; pinned host memory
// lets one slot's copies overlap the others' sorts).  The caller must not
// touch a request's buffers until it is done.  Results are exactly those of
// gs_sort_keys / gs_sort_pairs on the same input.
//
// serial = 1 turns the same server into the one-call-per-request baseline:
// one slot, no coalescing, one request at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gpusort_hip.h"
#include "gs_capi_impl.h"
#include "gs_internal.h"

using namespace gs;

namespace {

struct CopyJob {
  const void* src;
  void* dst;
  uint64_t count;  // elements
};

// One workgroup per job: dst[i] = src[i] for i < count.
template <typename T>
__global__ __launch_bounds__(256) void copy_jobs_kernel(const CopyJob* __restrict__ jobs) {
  const CopyJob j = jobs[blockIdx.x];
  const T* __restrict__ s = reinterpret_cast<const T*>(j.src);
  T* __restrict__ d = reinterpret_cast<T*>(j.dst);
  for (uint64_t i = threadIdx.x; i < j.count; i += 256) d[i] = s[i];
}

hipError_t copy_jobs(int elem_bytes, const CopyJob* d_jobs, uint32_t njobs, hipStream_t stream) {
  if (njobs == 0) return hipSuccess;
  if (elem_bytes == 4)
    hipLaunchKernelGGL(copy_jobs_kernel<uint32_t>, dim3(njobs), dim3(256), 0, stream, d_jobs);
  else
    hipLaunchKernelGGL(copy_jobs_kernel<uint64_t>, dim3(njobs), dim3(256), 0, stream, d_jobs);
  return hipGetLastError();
}

struct Request {
  uint64_t id = 0;
  gs_dtype kdt = GS_U32, vdt = GS_NONE;
  void* keys = nullptr;  // device memory, or host memory when `host`
  void* vals = nullptr;
  bool host = false;
  uint64_t n = 0;
  int desc = 0;
  double submit_us = 0, start_us = -1, done_us = -1;
  double dev_start_us = -1, dev_done_us = -1;
  gs_status status = GS_OK;
  uint64_t batch = 0;
  uint32_t batch_size = 0;
  bool done = false;
};

struct Slot {
  gs_handle* h = nullptr;
  hipStream_t stream = nullptr;  // the slot's own non-blocking stream (also h->stream)
  hipEvent_t ev = nullptr;       // batch completion
  hipEvent_t ev_start = nullptr; // batch start
  bool busy = false;
  std::vector<uint64_t> reqs;
  std::vector<unsigned char> meta;  // host sources of in-flight uploads
  std::vector<CopyJob> jobs;
};

double now_us(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct gs_server {
  gs_server_config cfg{};
  std::vector<Slot> slots;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request> reqs;  // indexed by id; push_back keeps references valid
  std::deque<uint64_t> pending;
  bool stop = false;
  gs_server_stats st{};
  std::chrono::steady_clock::time_point t0;
  hipEvent_t ev0 = nullptr;  // device-clock origin of dev_start_us / dev_done_us
  uint64_t next_batch = 1;
  // The one slot whose sorts may hand out tiles grouped by XCD (-1: none).
  // Grouped hand-out relies on the pass getting enough resident workgroups,
  // which holds when no other grouped pass runs beside it; every other slot
  // uses the arrival-order ticket, which progresses whatever else runs.
  int grouped_slot = -1;

  size_t pooled_bytes() const {
    size_t b = 0;
    for (const Slot& s : slots) b += s.h->arena.bytes();
    return b;
  }
  void run();
  bool admit_one(Slot& slot, std::unique_lock<std::mutex>& lk);
  gs_status launch(Slot& slot, const std::vector<Request>& batch);
};

namespace {

// Pool bytes a batch may add to a slot, an upper bound that follows the
// arena's growth rule slot by slot: one request sorted on its own, or
// `count` requests of `total` keys coalesced.
size_t need_single(gs_dtype kdt, gs_dtype vdt, uint64_t n, bool host) {
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  size_t b = n <= block_sort_max(kb, vb) ? 0 : Arena::grown(radix_workspace_bytes(kb, vb, n));
  if (host) b += Arena::grown((size_t)n * kb) + (vb ? Arena::grown((size_t)n * vb) : 0);  // device staging
  return b;
}

size_t need_coalesced(gs_dtype kdt, gs_dtype vdt, uint64_t total, uint64_t count) {
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  // Upper bound for any tile geometry: the smallest tiles are geometry 1's.
  const uint64_t tiles = total / radix_seg_tile_keys(kb, vb, 1) + count;
  size_t b = Arena::grown((size_t)total * kb) + Arena::grown(4 * count * sizeof(CopyJob)) +
             Arena::grown(radix_segments_workspace_bytes(kb, vb, total, tiles, (uint32_t)count)) +
             Arena::grown(tiles * 4 + count * 32 + (count + 1) * 8 + 4 * 256);
  if (vb) b += Arena::grown((size_t)total * vb);
  return b;
}

}  // namespace

// Queues one batch on the slot's stream (without the lock: `batch` holds
// copies of the requests).
gs_status gs_server::launch(Slot& slot, const std::vector<Request>& batch) {
  gs_handle* h = slot.h;
  const Request& r0 = batch[0];
  DeviceScope ds(h->device);
  {
    const hipError_t e = hipEventRecord(slot.ev_start, h->stream);
    if (e != hipSuccess) return from_hip(e, "record start");
  }
  if (batch.size() == 1 && !r0.host) {
    gs_status s = sort_device(h, r0.kdt, r0.keys, r0.vdt, r0.vals, r0.n, r0.desc);
    if (s != GS_OK) return s;
    return from_hip(hipEventRecord(slot.ev, h->stream), "record completion");
  }
  if (r0.host) {
    // Host requests: copied into the slot's staging buffers (back to back for
    // a coalesced batch), sorted there, copied back.  With pinned host
    // memory the copies of one slot overlap the sorts of the others.
    const int kb = dtype_bytes(r0.kdt), vb = val_bytes_of(r0.vdt);
    const uint32_t nreq = (uint32_t)batch.size();
    std::vector<uint64_t> off(nreq + 1, 0);
    for (uint32_t i = 0; i < nreq; ++i) off[i + 1] = off[i] + batch[i].n;
    void *sk = nullptr, *sv = nullptr;
    hipError_t e = h->arena.get(Arena::kStageKeys, off[nreq] * kb, h->stream, &sk);
    if (e == hipSuccess && vb) e = h->arena.get(Arena::kStageVals, off[nreq] * vb, h->stream, &sv);
    if (e != hipSuccess) return from_hip(e, "server staging");
    char* ck = reinterpret_cast<char*>(sk);
    char* cv = reinterpret_cast<char*>(sv);
    for (uint32_t i = 0; i < nreq && e == hipSuccess; ++i) {
      e = hipMemcpyAsync(ck + off[i] * kb, batch[i].keys, batch[i].n * kb, hipMemcpyHostToDevice, h->stream);
      if (e == hipSuccess && vb)
        e = hipMemcpyAsync(cv + off[i] * vb, batch[i].vals, batch[i].n * vb, hipMemcpyHostToDevice, h->stream);
    }
    if (e != hipSuccess) return from_hip(e, "server copy in");
    gs_status s = GS_OK;
    if (nreq == 1) {
      s = sort_device(h, r0.kdt, sk, r0.vdt, sv, r0.n, r0.desc);
    } else {
      std::vector<uint32_t> huge;
      s = sort_segments_async(h, r0.kdt, sk, r0.vdt, sv, off.data(), nreq, r0.desc, &slot.meta, &huge);
    }
    if (s != GS_OK) return s;
    for (uint32_t i = 0; i < nreq && e == hipSuccess; ++i) {
      e = hipMemcpyAsync(batch[i].keys, ck + off[i] * kb, batch[i].n * kb, hipMemcpyDeviceToHost, h->stream);
      if (e == hipSuccess && vb)
        e = hipMemcpyAsync(batch[i].vals, cv + off[i] * vb, batch[i].n * vb, hipMemcpyDeviceToHost, h->stream);
    }
    if (e != hipSuccess) return from_hip(e, "server copy out");
    return from_hip(hipEventRecord(slot.ev, h->stream), "record completion");
  }
  // Coalesced: gather -> segmented sort -> scatter.
  const int kb = dtype_bytes(r0.kdt), vb = val_bytes_of(r0.vdt);
  const uint32_t nreq = (uint32_t)batch.size();
  std::vector<uint64_t> off(nreq + 1, 0);
  for (uint32_t i = 0; i < nreq; ++i) off[i + 1] = off[i] + batch[i].n;
  const uint64_t total = off[nreq];
  void *sk = nullptr, *sv = nullptr, *dj = nullptr;
  hipError_t e = h->arena.get(Arena::kStageKeys, total * kb, h->stream, &sk);
  if (e == hipSuccess && vb) e = h->arena.get(Arena::kStageVals, total * vb, h->stream, &sv);
  const uint32_t groups = vb ? 4 : 2;  // gather keys, gather vals, scatter keys, scatter vals
  if (e == hipSuccess) e = h->arena.get(Arena::kPipeAux, (size_t)groups * nreq * sizeof(CopyJob), h->stream, &dj);
  if (e != hipSuccess) return from_hip(e, "server staging");
  slot.jobs.assign((size_t)groups * nreq, CopyJob{});
  for (uint32_t i = 0; i < nreq; ++i) {
    const Request& r = batch[i];
    char* kst = reinterpret_cast<char*>(sk) + off[i] * kb;
    slot.jobs[i] = CopyJob{r.keys, kst, r.n};
    slot.jobs[(vb ? 2 : 1) * nreq + i] = CopyJob{kst, r.keys, r.n};
    if (vb) {
      char* vst = reinterpret_cast<char*>(sv) + off[i] * vb;
      slot.jobs[nreq + i] = CopyJob{r.vals, vst, r.n};
      slot.jobs[3 * nreq + i] = CopyJob{vst, r.vals, r.n};
    }
  }
  e = hipMemcpyAsync(dj, slot.jobs.data(), slot.jobs.size() * sizeof(CopyJob), hipMemcpyHostToDevice, h->stream);
  const CopyJob* jobs = reinterpret_cast<const CopyJob*>(dj);
  if (e == hipSuccess) e = copy_jobs(kb, jobs, nreq, h->stream);
  if (e == hipSuccess && vb) e = copy_jobs(vb, jobs + nreq, nreq, h->stream);
  if (e != hipSuccess) return from_hip(e, "server gather");
  std::vector<uint32_t> huge;
  gs_status s = sort_segments_async(h, r0.kdt, sk, r0.vdt, sv, off.data(), nreq, r0.desc, &slot.meta, &huge);
  if (s != GS_OK) return s;
  e = copy_jobs(kb, jobs + (vb ? 2 : 1) * nreq, nreq, h->stream);
  if (e == hipSuccess && vb) e = copy_jobs(vb, jobs + 3 * nreq, nreq, h->stream);
  if (e != hipSuccess) return from_hip(e, "server scatter");
  return from_hip(hipEventRecord(slot.ev, h->stream), "record completion");
}

// Picks the next batch for a free slot; called with the lock held, releases
// it around the launch.  Returns true when something was launched.
bool gs_server::admit_one(Slot& slot, std::unique_lock<std::mutex>& lk) {
  if (pending.empty()) return false;
  const Request& head = reqs[pending.front()];
  std::vector<uint64_t> ids;
  bool coalesced = false;
  size_t need = 0;
  if (!cfg.serial && head.n <= cfg.coalesce_max_keys) {
    // Same-kind small requests in arrival order, while the batch's workspace
    // stays within the whole budget (so it can always be admitted once the
    // other slots drain).
    coalesced = true;
    uint64_t keys = 0;
    for (uint64_t id : pending) {
      const Request& r = reqs[id];
      if (r.n > cfg.coalesce_max_keys || r.kdt != head.kdt || r.vdt != head.vdt || r.desc != head.desc ||
          r.host != head.host)
        continue;
      const size_t trial = need_coalesced(head.kdt, head.vdt, keys + r.n, ids.size() + 1);
      if (!ids.empty() && (keys + r.n > cfg.batch_max_keys || ids.size() >= cfg.batch_max_requests ||
                           trial > cfg.hbm_budget))
        break;
      ids.push_back(id);
      keys += r.n;
      need = trial;
    }
  } else {
    ids.push_back(pending.front());
    need = need_single(head.kdt, head.vdt, head.n, head.host);
  }
  // Memory: the slot may grow by `need`; idle pools are released first when
  // the budget would be exceeded; otherwise the batch waits for busy slots.
  if (pooled_bytes() + need > cfg.hbm_budget) {
    slot.h->arena.release();
    for (Slot& o : slots)
      if (!o.busy && pooled_bytes() + need > cfg.hbm_budget) o.h->arena.release();
    if (pooled_bytes() + need > cfg.hbm_budget) return false;
  }
  for (uint64_t id : ids) pending.erase(std::find(pending.begin(), pending.end(), id));
  const uint64_t batch = next_batch++;
  const double t = now_us(t0);
  for (uint64_t id : ids) {
    reqs[id].start_us = t;
    reqs[id].batch = batch;
    reqs[id].batch_size = (uint32_t)ids.size();
  }
  slot.busy = true;
  slot.reqs = ids;
  {
    const int me = (int)(&slot - slots.data());
    if (grouped_slot < 0) grouped_slot = me;
    slot.h->tile_policy = grouped_slot == me ? kTilesGrouped : kTilesArrival;
  }
  std::vector<Request> copies;
  for (uint64_t id : ids) copies.push_back(reqs[id]);
  st.batches++;
  if (coalesced) {
    st.coalesced_batches++;
    st.coalesced_requests += ids.size();
  }
  lk.unlock();
  const gs_status s = launch(slot, copies);
  lk.lock();
  if (s != GS_OK) {
    const double td = now_us(t0);
    for (uint64_t id : ids) {
      reqs[id].status = s;
      reqs[id].done = true;
      reqs[id].done_us = td;
    }
    st.failed += ids.size();
    st.completed += ids.size();
    if (grouped_slot == (int)(&slot - slots.data())) grouped_slot = -1;
    slot.busy = false;
    cv_done.notify_all();
  }
  st.pooled_bytes = pooled_bytes();
  if (st.pooled_bytes > st.peak_pooled_bytes) st.peak_pooled_bytes = st.pooled_bytes;
  return true;
}

void gs_server::run() {
  hipSetDevice(cfg.device);
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    bool progress = false;
    // Completions.
    for (Slot& s : slots) {
      if (!s.busy) continue;
      lk.unlock();
      const hipError_t q = hipEventQuery(s.ev);
      lk.lock();
      if (q == hipErrorNotReady) continue;
      const double t = now_us(t0);
      float ms_start = -1.f, ms_done = -1.f;
      if (q == hipSuccess) {
        if (hipEventElapsedTime(&ms_start, ev0, s.ev_start) != hipSuccess) ms_start = -1.f;
        if (hipEventElapsedTime(&ms_done, ev0, s.ev) != hipSuccess) ms_done = -1.f;
      }
      for (uint64_t id : s.reqs) {
        reqs[id].done = true;
        reqs[id].done_us = t;
        reqs[id].dev_start_us = ms_start < 0 ? -1.0 : 1e3 * ms_start;
        reqs[id].dev_done_us = ms_done < 0 ? -1.0 : 1e3 * ms_done;
        if (q != hipSuccess) reqs[id].status = GS_DEVICE_ERROR;
      }
      st.completed += s.reqs.size();
      if (q != hipSuccess) st.failed += s.reqs.size();
      if (grouped_slot == (int)(&s - slots.data())) grouped_slot = -1;
      s.busy = false;
      s.reqs.clear();
      progress = true;
      cv_done.notify_all();
    }
    // Admission, in arrival order, one batch per free slot.
    for (Slot& s : slots) {
      if (s.busy || pending.empty()) continue;
      if (!admit_one(s, lk)) break;
      progress = true;
    }
    bool busy = false;
    for (const Slot& s : slots) busy = busy || s.busy;
    if (stop && !busy && pending.empty()) break;
    if (!progress) {
      if (busy)
        cv_work.wait_for(lk, std::chrono::microseconds(20));
      else
        cv_work.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
}

extern "C" {

gs_status gs_server_create(gs_server** out, const gs_server_config* cfg) {
  if (!out || !cfg) return fail(GS_INVALID_ARG, "null argument");
  *out = nullptr;
  if (cfg->streams == 0 || cfg->streams > 8) return fail(GS_INVALID_ARG, "streams must be 1..8");
  if (cfg->hbm_budget == 0) return fail(GS_INVALID_ARG, "zero HBM budget");
  gs_server* s = new gs_server();
  s->cfg = *cfg;
  if (s->cfg.serial) {
    s->cfg.streams = 1;
    s->cfg.coalesce_max_keys = 0;
  }
  if (s->cfg.batch_max_keys == 0) s->cfg.batch_max_keys = 1ull << 24;
  if (s->cfg.batch_max_requests == 0) s->cfg.batch_max_requests = 4096;
  s->t0 = std::chrono::steady_clock::now();
  s->st.budget = s->cfg.hbm_budget;
  auto teardown = [s]() {
    DeviceScope ds(s->cfg.device);
    for (Slot& o : s->slots) {
      gs_destroy(o.h);  // waits for the slot's stream
      if (o.ev) (void)hipEventDestroy(o.ev);
      if (o.ev_start) (void)hipEventDestroy(o.ev_start);
      if (o.stream) (void)hipStreamDestroy(o.stream);
    }
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    delete s;
  };
  for (uint32_t i = 0; i < s->cfg.streams; ++i) {
    Slot slot;
    gs_status st = gs_create(&slot.h, cfg->device);
    if (st == GS_OK) {
      DeviceScope ds(cfg->device);
      hipError_t e = hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreate(&slot.ev);
      if (e == hipSuccess) e = hipEventCreate(&slot.ev_start);
      if (e == hipSuccess) slot.h->stream = slot.stream;
      if (e != hipSuccess) st = from_hip(e, "server slot stream and events");
    }
    s->slots.push_back(slot);
    if (st != GS_OK) {
      teardown();
      return st;
    }
  }
  {
    DeviceScope ds(cfg->device);
    hipError_t e = hipEventCreate(&s->ev0);
    if (e == hipSuccess) e = hipEventRecord(s->ev0, s->slots[0].stream);
    if (e == hipSuccess) e = hipEventSynchronize(s->ev0);
    if (e != hipSuccess) {
      const gs_status st = from_hip(e, "server clock event");
      teardown();
      return st;
    }
  }
  s->th = std::thread([s] { s->run(); });
  *out = s;
  return GS_OK;
}

static gs_status server_submit(gs_server* s, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals, uint64_t n,
                               int descending, bool host, uint64_t* id) {
  if (!s || !id) return fail(GS_INVALID_ARG, "null argument");
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (vdt != GS_NONE && !dtype_ok(vdt)) return fail(GS_INVALID_ARG, "unsupported value type %d", (int)vdt);
  if (n > 0 && (!d_keys || (vdt != GS_NONE && !d_vals))) return fail(GS_INVALID_ARG, "null key or value pointer");
  Request r;
  r.kdt = kdt;
  r.vdt = vdt;
  r.keys = d_keys;
  r.vals = vdt == GS_NONE ? nullptr : d_vals;
  r.n = n;
  r.desc = descending != 0;
  r.host = host;
  {
    const bool small = !s->cfg.serial && n <= s->cfg.coalesce_max_keys;
    const size_t need = small ? need_coalesced(kdt, vdt, n, 1) : need_single(kdt, vdt, n, host);
    if (need > s->cfg.hbm_budget) {
      std::lock_guard<std::mutex> lk(s->mu);
      s->st.refused++;
      return fail(GS_ADMISSION_REFUSED, "request of %llu keys needs more workspace than the %llu-byte budget",
                  (unsigned long long)n, (unsigned long long)s->cfg.hbm_budget);
    }
  }
  std::lock_guard<std::mutex> lk(s->mu);
  r.id = s->reqs.size();
  r.submit_us = now_us(s->t0);
  if (n <= 1) {  // nothing to sort
    r.start_us = r.done_us = r.submit_us;
    r.done = true;
    s->st.completed++;
  } else {
    s->pending.push_back(r.id);
  }
  s->reqs.push_back(r);
  s->st.submitted++;
  *id = r.id;
  s->cv_work.notify_one();
  return GS_OK;
}

gs_status gs_server_submit(gs_server* s, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals, uint64_t n,
                           int descending, uint64_t* id) {
  return server_submit(s, kdt, d_keys, vdt, d_vals, n, descending, false, id);
}

gs_status gs_server_submit_host(gs_server* s, gs_dtype kdt, void* h_keys, gs_dtype vdt, void* h_vals, uint64_t n,
                                int descending, uint64_t* id) {
  return server_submit(s, kdt, h_keys, vdt, h_vals, n, descending, true, id);
}

gs_status gs_server_wait(gs_server* s, uint64_t id, int64_t timeout_us) {
  if (!s) return fail(GS_INVALID_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  if (id >= s->reqs.size()) return fail(GS_INVALID_ARG, "unknown request %llu", (unsigned long long)id);
  auto done = [&] { return s->reqs[id].done; };
  if (timeout_us < 0)
    s->cv_done.wait(lk, done);
  else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), done))
    return GS_NOT_READY;
  const gs_status st = s->reqs[id].status;
  return st == GS_OK ? GS_OK : fail(st, "request %llu failed on the device", (unsigned long long)id);
}

gs_status gs_server_drain(gs_server* s) {
  if (!s) return fail(GS_INVALID_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->st.completed == s->st.submitted; });
  return s->st.failed ? fail(GS_DEVICE_ERROR, "%llu request(s) failed", (unsigned long long)s->st.failed) : GS_OK;
}

gs_status gs_server_info(gs_server* s, uint64_t id, gs_request_info* out) {
  if (!s || !out) return fail(GS_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (id >= s->reqs.size()) return fail(GS_INVALID_ARG, "unknown request %llu", (unsigned long long)id);
  const Request& r = s->reqs[id];
  out->id = r.id;
  out->n = r.n;
  out->kdt = r.kdt;
  out->vdt = r.vdt;
  out->status = r.status;
  out->done = r.done ? 1 : 0;
  out->submit_us = r.submit_us;
  out->start_us = r.start_us;
  out->done_us = r.done_us;
  out->batch = r.batch;
  out->batch_size = r.batch_size;
  out->dev_start_us = r.dev_start_us;
  out->dev_done_us = r.dev_done_us;
  return GS_OK;
}

gs_status gs_server_get_stats(gs_server* s, gs_server_stats* out) {
  if (!s || !out) return fail(GS_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  return GS_OK;
}

gs_status gs_server_destroy(gs_server* s) {
  if (!s) return GS_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  s->th.join();
  DeviceScope ds(s->cfg.device);
  for (Slot& o : s->slots) {
    gs_destroy(o.h);  // waits for the slot's stream
    (void)hipEventDestroy(o.ev);
    (void)hipEventDestroy(o.ev_start);
    (void)hipStreamDestroy(o.stream);
  }
  (void)hipEventDestroy(s->ev0);
  delete s;
  return GS_OK;
}

}  // extern "C"
