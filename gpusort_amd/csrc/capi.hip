// C ABI of the sort path (declared in include/gpusort_hip.h).
//
// Every entry validates its arguments, switches to the handle's device,
// enqueues kernels on the handle's stream and maps HIP errors to gs_status
// plus a thread-local message.  No C++ exception crosses this boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpusort_hip.h"
#include "gs_capi_impl.h"
#include "gs_internal.h"

namespace gs {

static thread_local std::string g_last_error;

gs_status fail(gs_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

gs_status ok() { return GS_OK; }

gs_status from_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return GS_OK;
  if (e == hipErrorOutOfMemory) return fail(GS_OUT_OF_MEMORY, "%s: out of device memory", what);
  if (e == hipErrorNoDevice) return fail(GS_NO_DEVICE, "%s: no HIP device", what);
  if (e == hipErrorInvalidValue) return fail(GS_INVALID_ARG, "%s: invalid argument", what);
  return fail(GS_DEVICE_ERROR, "%s: %s", what, hipGetErrorString(e));
}

// ---------------------------------------------------------------- arena
Arena::~Arena() { release(); }

hipError_t Arena::get(int slot, size_t bytes, hipStream_t stream, void** out) {
  if (slot < 0 || slot >= kSlots) return hipErrorInvalidValue;
  if (bytes == 0) bytes = 256;
  if (cap_[slot] < bytes) {
    // Work queued on the stream may still use the old block.
    if (ptr_[slot]) {
      hipError_t e = hipStreamSynchronize(stream);
      if (e != hipSuccess) return e;
      e = hipFree(ptr_[slot]);
      if (e != hipSuccess) return e;
      ptr_[slot] = nullptr;
      cap_[slot] = 0;
    }
    const size_t want = grown(bytes);
    hipError_t e = hipMalloc(&ptr_[slot], want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      e = hipMalloc(&ptr_[slot], bytes);
      if (e != hipSuccess) {
        ptr_[slot] = nullptr;
        (void)hipGetLastError();
        return hipErrorOutOfMemory;
      }
      cap_[slot] = bytes;
    } else {
      cap_[slot] = want;
    }
  }
  *out = ptr_[slot];
  return hipSuccess;
}

size_t Arena::bytes() const {
  size_t s = 0;
  for (int i = 0; i < kSlots; ++i) s += cap_[i];
  return s;
}

void Arena::release() {
  for (int i = 0; i < kSlots; ++i) {
    if (ptr_[i]) (void)hipFree(ptr_[i]);
    ptr_[i] = nullptr;
    cap_[i] = 0;
  }
}

// ---------------------------------------------------------------- profiler
Profiler::~Profiler() {
  for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t Profiler::take() {
  if (used_ == pool_.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    pool_.push_back(e);
  }
  return pool_[used_++];
}

void Profiler::begin(const char* name, uint64_t bytes, hipStream_t stream) {
  Rec r{name, bytes, take(), nullptr};
  if (r.start) (void)hipEventRecord(r.start, stream);
  recs.push_back(r);
}

void Profiler::end(hipStream_t stream) {
  if (recs.empty()) return;
  Rec& r = recs.back();
  r.stop = take();
  if (r.stop) (void)hipEventRecord(r.stop, stream);
}

void Profiler::reset() {
  recs.clear();
  used_ = 0;
}

DeviceScope::DeviceScope(int dev) {
  if (hipGetDevice(&prev_) != hipSuccess) prev_ = -1;
  if (prev_ != dev) (void)hipSetDevice(dev);
  dev_ = dev;
}
DeviceScope::~DeviceScope() {
  if (prev_ >= 0 && prev_ != dev_) (void)hipSetDevice(prev_);
}

bool dtype_ok(gs_dtype d) { return d >= GS_I32 && d <= GS_F64; }
int dtype_bytes(gs_dtype d) { return (d == GS_I32 || d == GS_U32 || d == GS_F32) ? 4 : 8; }
int val_bytes_of(gs_dtype d) { return d == GS_NONE ? 0 : dtype_bytes(d); }

// Device sort of one array, small inputs in one workgroup, the rest onesweep.
gs_status sort_device(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals, uint64_t n,
                      int descending) {
  const int kb = dtype_bytes(kdt);
  const int vb = val_bytes_of(vdt);
  const XformParams xf = make_xform((int)kdt, descending != 0);
  if (n <= 1) return GS_OK;
  if (n <= block_sort_max(kb, vb)) {
    return from_hip(block_sort(kb, vb, d_keys, d_vals, (uint32_t)n, xf, h->stream), "block sort");
  }
  const size_t ws_bytes = radix_workspace_bytes(kb, vb, n);
  void* ws = nullptr;
  hipError_t e = h->arena.get(Arena::kRadix, ws_bytes, h->stream, &ws);
  if (e != hipSuccess) return from_hip(e, "radix workspace");
  return from_hip(radix_sort(kb, vb, d_keys, d_vals, n, xf, ws, ws_bytes, h->stream, h->timer(), h->tile_policy),
                  "radix sort");
}

}  // namespace gs

using namespace gs;

#define GS_CHECK_HANDLE(h) \
  if (!(h)) return fail(GS_INVALID_ARG, "null handle")

extern "C" {

const char* gs_version(void) { return "gpusort-mi355x 0.1 (gfx950)"; }

const char* gs_last_error(void) { return g_last_error.c_str(); }

gs_status gs_get_device_count(int* count) {
  if (!count) return fail(GS_INVALID_ARG, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e == hipErrorNoDevice || (e == hipSuccess && n == 0)) {
    *count = 0;
    (void)hipGetLastError();
    return GS_OK;
  }
  if (e != hipSuccess) {
    *count = 0;
    return from_hip(e, "hipGetDeviceCount");
  }
  *count = n;
  return GS_OK;
}

gs_status gs_create(gs_handle** out, int device) {
  if (!out) return fail(GS_INVALID_ARG, "null output handle");
  *out = nullptr;
  int n = 0;
  gs_status st = gs_get_device_count(&n);
  if (st != GS_OK) return st;
  if (n == 0) return fail(GS_NO_DEVICE, "No HIP-capable devices were detected");
  if (device < 0 || device >= n)
    return fail(GS_INVALID_ARG, "The device ordinal %d does not correspond to a valid device", device);
  gs_handle* h = new (std::nothrow) gs_handle();
  if (!h) return fail(GS_OUT_OF_MEMORY, "host allocation failed");
  h->device = device;
  h->stream = nullptr;
  DeviceScope ds(device);
  hipError_t e = hipFree(nullptr);  // initialise the context on this device
  if (e == hipSuccess) e = noalloc_setup(device);
  if (e != hipSuccess) {
    delete h;
    return from_hip(e, "device init");
  }
  *out = h;
  return GS_OK;
}

gs_status gs_destroy(gs_handle* h) {
  if (!h) return GS_OK;
  {
    DeviceScope ds(h->device);
    (void)hipStreamSynchronize(h->stream);
    h->arena.release();
    if (h->pinned) (void)hipHostFree(h->pinned);
  }
  delete h;
  return GS_OK;
}

gs_status gs_set_stream(gs_handle* h, void* stream) {
  GS_CHECK_HANDLE(h);
  h->stream = reinterpret_cast<hipStream_t>(stream);
  return GS_OK;
}

gs_status gs_set_tile_order(gs_handle* h, int order) {
  GS_CHECK_HANDLE(h);
  if (order != GS_TILES_AUTO && order != GS_TILES_ARRIVAL && order != GS_TILES_GROUPED)
    return fail(GS_INVALID_ARG, "gs_set_tile_order: unknown order %d", order);
  static_assert(GS_TILES_AUTO == kTilesByStream && GS_TILES_ARRIVAL == kTilesArrival &&
                    GS_TILES_GROUPED == kTilesGrouped,
                "the C ABI's tile orders are the library's tile policies");
  h->tile_policy = order;
  return GS_OK;
}

gs_status gs_synchronize(gs_handle* h) {
  GS_CHECK_HANDLE(h);
  DeviceScope ds(h->device);
  return from_hip(hipStreamSynchronize(h->stream), "synchronize");
}

gs_status gs_arena_bytes(gs_handle* h, uint64_t* bytes) {
  GS_CHECK_HANDLE(h);
  if (!bytes) return fail(GS_INVALID_ARG, "null output");
  *bytes = h->arena.bytes();
  return GS_OK;
}

gs_status gs_release_arena(gs_handle* h) {
  GS_CHECK_HANDLE(h);
  DeviceScope ds(h->device);
  (void)hipStreamSynchronize(h->stream);
  h->arena.release();
  return GS_OK;
}

gs_status gs_sort_keys(gs_handle* h, gs_dtype kdt, void* d_keys, uint64_t n, int descending) {
  GS_CHECK_HANDLE(h);
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (n > 0 && !d_keys) return fail(GS_INVALID_ARG, "null key pointer");
  DeviceScope ds(h->device);
  return sort_device(h, kdt, d_keys, GS_NONE, nullptr, n, descending);
}

gs_status gs_sort_pairs(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals, uint64_t n,
                        int descending) {
  GS_CHECK_HANDLE(h);
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (!dtype_ok(vdt)) return fail(GS_INVALID_ARG, "unsupported value type %d", (int)vdt);
  if (n > 0 && (!d_keys || !d_vals)) return fail(GS_INVALID_ARG, "null key or value pointer");
  DeviceScope ds(h->device);
  return sort_device(h, kdt, d_keys, vdt, d_vals, n, descending);
}

gs_status gs_sort_keys_host(gs_handle* h, gs_dtype kdt, void* h_keys, uint64_t n, int descending) {
  GS_CHECK_HANDLE(h);
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (n == 0) return GS_OK;
  if (!h_keys) return fail(GS_INVALID_ARG, "null key pointer");
  DeviceScope ds(h->device);
  const size_t kbytes = n * (size_t)dtype_bytes(kdt);
  void* dk = nullptr;
  hipError_t e = h->arena.get(Arena::kStageKeys, kbytes, h->stream, &dk);
  if (e != hipSuccess) return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  e = hipMemcpyAsync(dk, h_keys, kbytes, hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) return from_hip(e, "copy keys to device");
  gs_status st = sort_device(h, kdt, dk, GS_NONE, nullptr, n, descending);
  if (st != GS_OK) return st;
  e = hipMemcpyAsync(h_keys, dk, kbytes, hipMemcpyDeviceToHost, h->stream);
  if (e != hipSuccess) return from_hip(e, "copy keys to host");
  return from_hip(hipStreamSynchronize(h->stream), "synchronize");
}

gs_status gs_sort_pairs_host(gs_handle* h, gs_dtype kdt, void* h_keys, gs_dtype vdt, void* h_vals, uint64_t n,
                             int descending, void* h_pairs_out, uint32_t pair_stride, uint32_t val_offset) {
  GS_CHECK_HANDLE(h);
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (!dtype_ok(vdt)) return fail(GS_INVALID_ARG, "unsupported value type %d", (int)vdt);
  if (n == 0) return GS_OK;
  if (!h_keys || !h_vals) return fail(GS_INVALID_ARG, "null key or value pointer");
  const int kb = dtype_bytes(kdt), vb = dtype_bytes(vdt);
  if (h_pairs_out && (pair_stride < (uint32_t)(kb + vb) || val_offset < (uint32_t)kb ||
                      val_offset + vb > pair_stride))
    return fail(GS_INVALID_ARG, "pair layout (stride %u, value offset %u) cannot hold the key and value",
                pair_stride, val_offset);
  DeviceScope ds(h->device);
  void *dk = nullptr, *dv = nullptr;
  if (h->arena.get(Arena::kStageKeys, n * (size_t)kb, h->stream, &dk) != hipSuccess ||
      h->arena.get(Arena::kStageVals, n * (size_t)vb, h->stream, &dv) != hipSuccess)
    return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  hipError_t e = hipMemcpyAsync(dk, h_keys, n * (size_t)kb, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dv, h_vals, n * (size_t)vb, hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) return from_hip(e, "copy to device");
  gs_status st = sort_device(h, kdt, dk, vdt, dv, n, descending);
  if (st != GS_OK) return st;
  e = hipMemcpyAsync(h_keys, dk, n * (size_t)kb, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h_vals, dv, n * (size_t)vb, hipMemcpyDeviceToHost, h->stream);
  if (e != hipSuccess) return from_hip(e, "copy to host");
  if (h_pairs_out) {
    void* dp = nullptr;
    if (h->arena.get(Arena::kStagePairs, n * (size_t)pair_stride, h->stream, &dp) != hipSuccess)
      return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
    e = pack_pairs(kb, dk, vb, dv, dp, n, pair_stride, val_offset, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h_pairs_out, dp, n * (size_t)pair_stride, hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return from_hip(e, "pack pairs");
  }
  return from_hip(hipStreamSynchronize(h->stream), "synchronize");
}

gs_status gs_multisplit(gs_handle* h, gs_dtype kdt, const void* d_keys, gs_dtype vdt, const void* d_vals, uint64_t n,
                        const void* h_splitters, uint32_t nspl, int upper, int descending, void* d_out_keys,
                        void* d_out_vals, uint64_t* h_bucket_offsets) {
  GS_CHECK_HANDLE(h);
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (vdt != GS_NONE && !dtype_ok(vdt)) return fail(GS_INVALID_ARG, "unsupported value type %d", (int)vdt);
  if (nspl > 255) return fail(GS_INVALID_ARG, "at most 255 splitters (256 buckets), got %u", nspl);
  if (!h_bucket_offsets) return fail(GS_INVALID_ARG, "null bucket offsets");
  if (nspl && !h_splitters) return fail(GS_INVALID_ARG, "null splitters");
  if (n > 0 && (!d_keys || !d_out_keys || (vdt != GS_NONE && (!d_vals || !d_out_vals))))
    return fail(GS_INVALID_ARG, "null key or value pointer");
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  if (n > 0 && (d_keys == d_out_keys || (vb && d_vals == d_out_vals)))
    return fail(GS_INVALID_ARG, "multisplit output must not alias its input");
  for (uint32_t j = 0; j <= nspl + 1; ++j) h_bucket_offsets[j] = 0;
  if (n == 0) return GS_OK;
  DeviceScope ds(h->device);
  const XformParams xf = make_xform((int)kdt, descending != 0);
  const size_t ws_bytes = multisplit_workspace_bytes(kb, vb, n);
  const size_t spl_off = align_up(ws_bytes, 256), base_off = spl_off + align_up((size_t)255 * 8, 256);
  void* ws = nullptr;
  hipError_t e = h->arena.get(Arena::kRadix, base_off + 256 * sizeof(uint64_t), h->stream, &ws);
  if (e != hipSuccess) return from_hip(e, "multisplit workspace");
  unsigned char* w = reinterpret_cast<unsigned char*>(ws);
  if (nspl) e = hipMemcpyAsync(w + spl_off, h_splitters, (size_t)nspl * kb, hipMemcpyHostToDevice, h->stream);
  uint64_t* d_base = reinterpret_cast<uint64_t*>(w + base_off);
  if (e == hipSuccess)
    e = multisplit(kb, vb, d_keys, d_vals, n, w + spl_off, nspl, upper ? 1 : 0, xf, d_out_keys, d_out_vals, d_base,
                   ws, ws_bytes, h->stream, h->timer(), h->tile_policy);
  uint64_t base[256];
  if (e == hipSuccess) e = hipMemcpyAsync(base, d_base, sizeof(base), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return from_hip(e, "multisplit");
  for (uint32_t j = 0; j <= nspl; ++j) h_bucket_offsets[j] = base[j];
  h_bucket_offsets[nspl + 1] = n;
  return GS_OK;
}

gs_status gs_merge(gs_handle* h, gs_dtype kdt, const void* a_k, const void* a_v, uint64_t na, const void* b_k,
                   const void* b_v, uint64_t nb, gs_dtype vdt, void* o_k, void* o_v, int descending) {
  GS_CHECK_HANDLE(h);
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (vdt != GS_NONE && !dtype_ok(vdt)) return fail(GS_INVALID_ARG, "unsupported value type %d", (int)vdt);
  const uint64_t total = na + nb;
  if (total == 0) return GS_OK;
  if ((na && !a_k) || (nb && !b_k) || !o_k) return fail(GS_INVALID_ARG, "null key pointer");
  if (vdt != GS_NONE && ((na && !a_v) || (nb && !b_v) || !o_v)) return fail(GS_INVALID_ARG, "null value pointer");
  DeviceScope ds(h->device);
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  const XformParams xf = make_xform((int)kdt, descending != 0);
  MergeJob job{0, na, 0, nb, 0};
  MergeJobDev dj[1];
  std::vector<uint32_t> tile_job;
  const uint32_t ntiles = merge_plan(&job, 1, dj, &tile_job);
  const size_t bytes = align_up(sizeof(MergeJobDev), 256) + align_up(ntiles * sizeof(uint32_t), 256) +
                       align_up((size_t)ntiles * sizeof(uint64_t), 256);
  void* ws = nullptr;
  hipError_t e = h->arena.get(Arena::kMerge, bytes, h->stream, &ws);
  if (e != hipSuccess) return from_hip(e, "merge workspace");
  char* w = reinterpret_cast<char*>(ws);
  MergeJobDev* d_jobs = reinterpret_cast<MergeJobDev*>(w);
  uint32_t* d_tile_job = reinterpret_cast<uint32_t*>(w + align_up(sizeof(MergeJobDev), 256));
  uint64_t* d_split = reinterpret_cast<uint64_t*>(w + align_up(sizeof(MergeJobDev), 256) +
                                                  align_up(ntiles * sizeof(uint32_t), 256));
  e = hipMemcpyAsync(d_jobs, dj, sizeof(MergeJobDev), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_tile_job, tile_job.data(), ntiles * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = merge_tiles(kb, vb, a_k, b_k, a_v, b_v, o_k, o_v, d_jobs, d_tile_job, ntiles, xf, d_split, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);  // the host job table is stack memory
  return from_hip(e, "merge");
}

gs_status gs_bucket_bounds(gs_handle* h, gs_dtype kdt, const void* d_keys, const uint64_t* h_run_off, uint32_t nruns,
                           const void* h_spl, uint32_t nspl, int upper, int descending, uint64_t* h_bounds) {
  GS_CHECK_HANDLE(h);
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (nruns == 0) return GS_OK;
  if (!h_run_off || !h_bounds || (nspl && !h_spl) || !d_keys) return fail(GS_INVALID_ARG, "null pointer");
  DeviceScope ds(h->device);
  const int kb = dtype_bytes(kdt);
  const size_t n_off = nruns + 1, n_b = (size_t)nruns * (nspl + 2);
  const size_t o_off = 0, o_spl = align_up(n_off * 8, 256), o_first = o_spl + align_up((size_t)nspl * kb + 8, 256),
               o_b = o_first + align_up((size_t)nruns * 4, 256), total = o_b + align_up(n_b * 8, 256);
  void* ws = nullptr;
  hipError_t e = h->arena.get(Arena::kSmall, total, h->stream, &ws);
  if (e != hipSuccess) return from_hip(e, "bounds workspace");
  char* w = reinterpret_cast<char*>(ws);
  std::vector<uint32_t> first(nruns, 0);
  e = hipMemcpyAsync(w + o_off, h_run_off, n_off * 8, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess && nspl)
    e = hipMemcpyAsync(w + o_spl, h_spl, (size_t)nspl * kb, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(w + o_first, first.data(), nruns * 4, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = bucket_bounds(kb, d_keys, reinterpret_cast<const uint64_t*>(w + o_off), nruns, w + o_spl,
                      reinterpret_cast<const uint32_t*>(w + o_first), nspl, upper,
                      make_xform((int)kdt, descending != 0), reinterpret_cast<uint64_t*>(w + o_b), h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h_bounds, w + o_b, n_b * 8, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return from_hip(e, "bucket bounds");
}

gs_status gs_profile_begin(gs_handle* h) {
  GS_CHECK_HANDLE(h);
  DeviceScope ds(h->device);
  h->profiler.reset();
  h->profiler.active = true;
  return GS_OK;
}

gs_status gs_profile_end(gs_handle* h, gs_kernel_stat* out, uint32_t max_out, uint32_t* n_out) {
  GS_CHECK_HANDLE(h);
  if (!n_out || (max_out && !out)) return fail(GS_INVALID_ARG, "null output");
  DeviceScope ds(h->device);
  h->profiler.active = false;
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return from_hip(e, "profile synchronize");
  std::vector<gs_kernel_stat> agg;
  for (const auto& r : h->profiler.recs) {
    float ms = 0.f;
    if (!r.start || !r.stop || hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess) continue;
    gs_kernel_stat* st = nullptr;
    for (auto& a : agg)
      if (r.name == a.name) st = &a;
    if (!st) {
      gs_kernel_stat z{};
      snprintf(z.name, sizeof(z.name), "%s", r.name.c_str());
      z.min_ms = 1e30;
      agg.push_back(z);
      st = &agg.back();
    }
    st->launches += 1;
    st->total_ms += ms;
    st->bytes += r.bytes;
    if (ms > st->max_ms) st->max_ms = ms;
    if (ms < st->min_ms) st->min_ms = ms;
  }
  h->profiler.reset();
  *n_out = (uint32_t)agg.size();
  for (uint32_t i = 0; i < agg.size() && i < max_out; ++i) out[i] = agg[i];
  return GS_OK;
}

gs_status gs_gather(gs_handle* h, int elem_bytes, const void* d_src, const uint64_t* d_idx, void* d_dst, uint64_t n) {
  GS_CHECK_HANDLE(h);
  if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
    return fail(GS_INVALID_ARG, "element size %d not in {1,2,4,8}", elem_bytes);
  if (n == 0) return GS_OK;
  if (!d_src || !d_idx || !d_dst) return fail(GS_INVALID_ARG, "null pointer");
  DeviceScope ds(h->device);
  LaunchTimer* t = h->timer();
  if (t) t->begin("gather", n * (8 + 2 * (uint64_t)elem_bytes), h->stream);
  const hipError_t e = gather_values(elem_bytes, d_src, d_idx, d_dst, n, h->stream);
  if (t) t->end(h->stream);
  return from_hip(e, "gather");
}

}  // extern "C"
