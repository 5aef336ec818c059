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
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/gpusort_hip.h"
#include "gs_capi_impl.h"
#include "gs_device.h"
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

// ------------------------------------------------------------ host chunks
//
// Host-memory sorts of inputs larger than the device can stage, as the
// reference's CudaUtils::Sort / SortByKey do (src/cuda_utils.cu:97-167,
// :169-243): chunks of budget / (3 * element bytes) elements
// (SplitChunks, :508-520, with the element = key + payload), each sorted on
// the device, then a binary tree of stable 2-way merges of neighbouring runs
// (ParallelMergeArrays, :323-371), ping-ponging between the caller's arrays
// and a host buffer so that the last level lands in the caller's arrays.
// Each 2-way merge streams through the device in output blocks of at most a
// chunk: the block's split of the two runs is a merge-path search over the
// ordered key images on the host, the two input slices go in, the device
// merge (ties from the earlier run) writes the block, and it comes back.

// Staging of at most this many bytes (3 * n * element bytes) takes one piece
// without asking the device for its free memory (the query costs about as
// much as a small sort's kernels); such a call that does not fit fails with
// GS_OUT_OF_MEMORY instead of being chunked.
constexpr uint64_t kHostQueryBytes = 256ull << 20;

// Elements one device-staged piece may hold for an n-element call; 0 when
// not even one fits.
static uint64_t host_chunk_elems(gs_handle* h, int kb, int vb, uint64_t n) {
  const uint64_t per = 3ull * (uint64_t)(kb + vb);
  uint64_t budget = h->host_budget;
  if (budget == 0) {
    if (n <= kHostQueryBytes / per) return n;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
    budget = free_b;
  }
  return budget / per;
}

// Number of elements of a (na) among the first d of the stable merge of a
// and b (ties from a), by a binary search over the ordered images.
template <typename U>
static uint64_t host_corank(const U* a, uint64_t na, const U* b, uint64_t nb, uint64_t d, const Xform<U>& xf) {
  uint64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (to_ordered(a[mid], xf) <= to_ordered(b[d - mid - 1], xf))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

static uint64_t corank_of(int kb, const void* a, uint64_t na, const void* b, uint64_t nb, uint64_t d,
                          const XformParams& xp) {
  if (kb == 4) {
    const Xform<uint32_t> x{(uint32_t)xp.float_mask, (uint32_t)xp.sign_flip, (uint32_t)xp.desc_mask};
    return host_corank<uint32_t>(static_cast<const uint32_t*>(a), na, static_cast<const uint32_t*>(b), nb, d, x);
  }
  const Xform<uint64_t> x{xp.float_mask, xp.sign_flip, xp.desc_mask};
  return host_corank<uint64_t>(static_cast<const uint64_t*>(a), na, static_cast<const uint64_t*>(b), nb, d, x);
}

// Stable merge of host runs a = keys[a0, a0 + na), b = the na.. that follow
// it in the same arrays (src), into dst at the same offsets, through the
// device in blocks of at most `block` elements.
static gs_status host_merge_runs(gs_handle* h, gs_dtype kdt, gs_dtype vdt, const unsigned char* src_k,
                                 const unsigned char* src_v, unsigned char* dst_k, unsigned char* dst_v, uint64_t a0,
                                 uint64_t na, uint64_t nb, uint64_t block, int descending) {
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  const XformParams xp = make_xform((int)kdt, descending != 0);
  const unsigned char* ak = src_k + a0 * kb;
  const unsigned char* bk = ak + na * kb;
  const unsigned char* av = vb ? src_v + a0 * vb : nullptr;
  const unsigned char* bv = vb ? av + na * vb : nullptr;
  void *din_k = nullptr, *din_v = nullptr, *dout_k = nullptr, *dout_v = nullptr;
  if (h->arena.get(Arena::kStageKeys, block * kb, h->stream, &din_k) != hipSuccess ||
      h->arena.get(Arena::kPipeKeys, block * kb, h->stream, &dout_k) != hipSuccess ||
      (vb && (h->arena.get(Arena::kStageVals, block * vb, h->stream, &din_v) != hipSuccess ||
              h->arena.get(Arena::kPipeVals, block * vb, h->stream, &dout_v) != hipSuccess)))
    return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  const uint64_t total = na + nb;
  uint64_t i0 = 0;
  for (uint64_t o = 0; o < total; o += block) {
    const uint64_t o2 = std::min(total, o + block);
    const uint64_t i1 = corank_of(kb, ak, na, bk, nb, o2, xp);
    const uint64_t j0 = o - i0, j1 = o2 - i1;
    const uint64_t ca = i1 - i0, cb = j1 - j0;
    unsigned char* ik = static_cast<unsigned char*>(din_k);
    unsigned char* iv = static_cast<unsigned char*>(din_v);
    hipError_t e = hipSuccess;
    if (ca) e = hipMemcpyAsync(ik, ak + i0 * kb, ca * kb, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess && cb) e = hipMemcpyAsync(ik + ca * kb, bk + j0 * kb, cb * kb, hipMemcpyHostToDevice, h->stream);
    if (vb) {
      if (e == hipSuccess && ca) e = hipMemcpyAsync(iv, av + i0 * vb, ca * vb, hipMemcpyHostToDevice, h->stream);
      if (e == hipSuccess && cb)
        e = hipMemcpyAsync(iv + ca * vb, bv + j0 * vb, cb * vb, hipMemcpyHostToDevice, h->stream);
    }
    if (e != hipSuccess) return from_hip(e, "copy merge block to device");
    const gs_status st = gs_merge(h, kdt, ik, vb ? iv : nullptr, ca, ik + ca * kb, vb ? iv + ca * vb : nullptr, cb,
                                  vb ? vdt : GS_NONE, dout_k, vb ? dout_v : nullptr, descending);
    if (st != GS_OK) return st;
    e = hipMemcpyAsync(dst_k + (a0 + o) * kb, dout_k, (o2 - o) * kb, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && vb)
      e = hipMemcpyAsync(dst_v + (a0 + o) * vb, dout_v, (o2 - o) * vb, hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return from_hip(e, "copy merge block to host");
    i0 = i1;
  }
  return from_hip(hipStreamSynchronize(h->stream), "synchronize");
}

// Packs sorted host keys and values into pair records through the device,
// `block` elements at a time.
static gs_status host_pack_pairs(gs_handle* h, int kb, const void* h_keys, int vb, const void* h_vals, uint64_t n,
                                 void* h_pairs_out, uint32_t stride, uint32_t val_offset, uint64_t block) {
  void *dk = nullptr, *dv = nullptr, *dp = nullptr;
  if (h->arena.get(Arena::kStageKeys, block * kb, h->stream, &dk) != hipSuccess ||
      h->arena.get(Arena::kStageVals, block * vb, h->stream, &dv) != hipSuccess ||
      h->arena.get(Arena::kStagePairs, block * stride, h->stream, &dp) != hipSuccess)
    return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  const unsigned char* hk = static_cast<const unsigned char*>(h_keys);
  const unsigned char* hv = static_cast<const unsigned char*>(h_vals);
  unsigned char* hp = static_cast<unsigned char*>(h_pairs_out);
  for (uint64_t s = 0; s < n; s += block) {
    const uint64_t m = std::min(block, n - s);
    hipError_t e = hipMemcpyAsync(dk, hk + s * kb, m * kb, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dv, hv + s * vb, m * vb, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = pack_pairs(kb, dk, vb, dv, dp, m, stride, val_offset, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hp + s * stride, dp, m * stride, hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return from_hip(e, "pack pairs");
  }
  return from_hip(hipStreamSynchronize(h->stream), "synchronize");
}

// Chunked sort of host keys (and payloads) in place; `chunk` < n.
static gs_status host_sort_chunked(gs_handle* h, gs_dtype kdt, void* h_keys, gs_dtype vdt, void* h_vals, uint64_t n,
                                   int descending, uint64_t chunk) {
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  const uint64_t nchunks = (n + chunk - 1) / chunk;
  int height = 0;
  while ((1ull << height) < nchunks) ++height;
  // Host buffer of the input's size, not zero-filled (every byte is written
  // before it is read).
  std::unique_ptr<unsigned char[]> tmp_k(new (std::nothrow) unsigned char[n * kb]);
  std::unique_ptr<unsigned char[]> tmp_v(vb ? new (std::nothrow) unsigned char[n * vb] : nullptr);
  if (!tmp_k || (vb && !tmp_v)) return fail(GS_OUT_OF_MEMORY, "host allocation failed");
  unsigned char* user_k = static_cast<unsigned char*>(h_keys);
  unsigned char* user_v = static_cast<unsigned char*>(h_vals);
  // The sorted chunks go where an even number of merge levels leaves the
  // result in the caller's arrays.
  unsigned char* cur_k = (height % 2) ? tmp_k.get() : user_k;
  unsigned char* cur_v = (height % 2) ? tmp_v.get() : user_v;
  unsigned char* oth_k = (height % 2) ? user_k : tmp_k.get();
  unsigned char* oth_v = (height % 2) ? user_v : tmp_v.get();
  void *dk = nullptr, *dv = nullptr;
  if (h->arena.get(Arena::kStageKeys, chunk * kb, h->stream, &dk) != hipSuccess ||
      (vb && h->arena.get(Arena::kStageVals, chunk * vb, h->stream, &dv) != hipSuccess))
    return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  for (uint64_t s = 0; s < n; s += chunk) {
    const uint64_t m = std::min(chunk, n - s);
    hipError_t e = hipMemcpyAsync(dk, user_k + s * kb, m * kb, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess && vb) e = hipMemcpyAsync(dv, user_v + s * vb, m * vb, hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) return from_hip(e, "copy chunk to device");
    const gs_status st = sort_device(h, kdt, dk, vdt, dv, m, descending);
    if (st != GS_OK) return st;
    e = hipMemcpyAsync(cur_k + s * kb, dk, m * kb, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && vb) e = hipMemcpyAsync(cur_v + s * vb, dv, m * vb, hipMemcpyDeviceToHost, h->stream);
    if (e != hipSuccess) return from_hip(e, "copy chunk to host");
  }
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return from_hip(e, "synchronize");
  // Merge levels: runs of `run` elements pair up; a run without a partner
  // is copied as it is.
  for (uint64_t run = chunk; run < n; run *= 2) {
    for (uint64_t a0 = 0; a0 < n; a0 += 2 * run) {
      const uint64_t na = std::min(run, n - a0);
      const uint64_t nb = std::min(run, n - a0 - na);
      if (nb == 0) {
        std::memcpy(oth_k + a0 * kb, cur_k + a0 * kb, na * kb);
        if (vb) std::memcpy(oth_v + a0 * vb, cur_v + a0 * vb, na * vb);
        continue;
      }
      const gs_status st = host_merge_runs(h, kdt, vdt, cur_k, cur_v, oth_k, oth_v, a0, na, nb, chunk, descending);
      if (st != GS_OK) return st;
    }
    std::swap(cur_k, oth_k);
    std::swap(cur_v, oth_v);
  }
  if (cur_k != user_k) {  // not reached: the chunks start where the level count puts the result
    std::memcpy(user_k, cur_k, n * kb);
    if (vb) std::memcpy(user_v, cur_v, n * vb);
  }
  return GS_OK;
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

gs_status gs_set_host_sort_budget(gs_handle* h, uint64_t bytes) {
  GS_CHECK_HANDLE(h);
  h->host_budget = bytes;
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
  const uint64_t chunk = host_chunk_elems(h, dtype_bytes(kdt), 0, n);
  if (chunk == 0) return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  if (chunk < n) return host_sort_chunked(h, kdt, h_keys, GS_NONE, nullptr, n, descending, chunk);
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
  const uint64_t chunk = host_chunk_elems(h, kb, vb, n);
  if (chunk == 0) return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  if (chunk < n) {
    gs_status st = host_sort_chunked(h, kdt, h_keys, vdt, h_vals, n, descending, chunk);
    if (st != GS_OK || !h_pairs_out) return st;
    return host_pack_pairs(h, kb, h_keys, vb, h_vals, n, h_pairs_out, pair_stride, val_offset, chunk);
  }
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
