// gs_sort_pairs_noalloc: the device sort with its workspace in a fixed device
// array of the library instead of the handle's arena, so a call allocates
// nothing (no hipMalloc / hipFree, no allocating library routine).
//
// The workspace of a radix sort of n keys (ping-pong copy of keys and
// payloads, digit histograms, look-back status words, control block) lives
// in `scratch`, one copy per GPU, loaded with this file's code object and
// held until the process ends.  Every call clears what must start at zero
// with kernels of its own (fill / histogram / scan launches), so the last
// call's contents never matter; two calls that overlap on one GPU share the
// array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gpusort_hip.h"
#include "gs_capi_impl.h"
#include "gs_internal.h"

namespace gs {

static __device__ __align__(256) unsigned char scratch[GS_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
// Device address of `scratch` on each GPU.  The array's name is a host-side
// stand-in and is never passed to a kernel or a copy.
void* g_scratch[kMaxDevices];
}  // namespace

// Resolves this GPU's copy of the array.  Called by gs_create with `device`
// current, once per device.
hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_scratch[device] = ptr;
  return hipSuccess;
}

static void* scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_scratch[device];
}

}  // namespace gs

using namespace gs;

extern "C" gs_status gs_sort_pairs_noalloc(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals,
                                           uint64_t n, int descending) {
  if (!h) return fail(GS_INVALID_ARG, "null handle");
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (vdt != GS_NONE && !dtype_ok(vdt)) return fail(GS_INVALID_ARG, "unsupported value type %d", (int)vdt);
  if (n > 0 && (!d_keys || (vdt != GS_NONE && !d_vals))) return fail(GS_INVALID_ARG, "null key or value pointer");
  if (n <= 1) return GS_OK;
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  DeviceScope ds(h->device);
  const XformParams xf = make_xform((int)kdt, descending != 0);
  if (n <= block_sort_max(kb, vb))  // one workgroup in LDS: no workspace
    return from_hip(block_sort(kb, vb, d_keys, d_vals, (uint32_t)n, xf, h->stream), "block sort");
  const size_t need = radix_workspace_bytes(kb, vb, n);
  if (need > (size_t)GS_NOALLOC_SCRATCH_BYTES)
    return fail(GS_INVALID_ARG,
                "gs_sort_pairs_noalloc: %llu keys of %d + %d bytes need %zu bytes of workspace, more than the "
                "%llu-byte fixed scratch array; use gs_sort_pairs",
                (unsigned long long)n, kb, vb, need, (unsigned long long)GS_NOALLOC_SCRATCH_BYTES);
  void* ws = scratch_of(h->device);
  if (!ws) return fail(GS_DEVICE_ERROR, "gs_sort_pairs_noalloc: the scratch array of device %d is not resolved",
                       h->device);
  return from_hip(radix_sort(kb, vb, d_keys, d_vals, n, xf, ws, (size_t)GS_NOALLOC_SCRATCH_BYTES, h->stream, nullptr,
                             h->tile_policy),
                  "radix sort");
}
