// Batched segmented sort (gs_sort_segments).
//
// Segments are binned by size class on the host from the offsets:
//   small  (<= one LDS tile): one workgroup per segment, bitonic in LDS, all
//          small segments in one launch
//   large  : onesweep radix sort per segment on the same stream
// Every class is stable with a payload, so each segment's result equals a
// separate gs_sort_pairs call on it.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/gpusort_hip.h"
#include "gs_capi_impl.h"
#include "gs_internal.h"

using namespace gs;

extern "C" gs_status gs_sort_segments(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals,
                                      const uint64_t* d_offsets, uint32_t nseg, int descending) {
  if (!h) return fail(GS_INVALID_ARG, "null handle");
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (vdt != GS_NONE && !dtype_ok(vdt)) return fail(GS_INVALID_ARG, "unsupported value type %d", (int)vdt);
  if (nseg == 0) return GS_OK;
  if (!d_offsets || !d_keys || (vdt != GS_NONE && !d_vals)) return fail(GS_INVALID_ARG, "null pointer");
  DeviceScope ds(h->device);
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  const XformParams xf = make_xform((int)kdt, descending != 0);
  std::vector<uint64_t> off(nseg + 1);
  hipError_t e = hipMemcpyAsync(off.data(), d_offsets, (nseg + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return from_hip(e, "read segment offsets");
  for (uint32_t s = 0; s < nseg; ++s)
    if (off[s + 1] < off[s]) return fail(GS_INVALID_ARG, "segment offsets decrease at segment %u", s);

  const uint32_t small_max = block_sort_max(kb, vb);
  std::vector<uint32_t> small_ids;
  std::vector<uint32_t> large_ids;
  for (uint32_t s = 0; s < nseg; ++s) {
    const uint64_t len = off[s + 1] - off[s];
    if (len <= 1) continue;
    if (len <= small_max)
      small_ids.push_back(s);
    else
      large_ids.push_back(s);
  }
  if (!small_ids.empty()) {
    void* ws = nullptr;
    e = h->arena.get(Arena::kSegAux, small_ids.size() * sizeof(uint32_t), h->stream, &ws);
    if (e != hipSuccess) return from_hip(e, "segment workspace");
    e = hipMemcpyAsync(ws, small_ids.data(), small_ids.size() * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
      e = block_sort_segments(kb, vb, d_keys, d_vals, d_offsets, reinterpret_cast<const uint32_t*>(ws),
                              (uint32_t)small_ids.size(), xf, h->stream);
    if (e != hipSuccess) return from_hip(e, "small segments");
  }
  char* kbase = reinterpret_cast<char*>(d_keys);
  char* vbase = reinterpret_cast<char*>(d_vals);
  for (uint32_t s : large_ids) {
    const uint64_t len = off[s + 1] - off[s];
    gs_status st = sort_device(h, kdt, kbase + off[s] * kb, vdt, vb ? vbase + off[s] * vb : nullptr, len, descending);
    if (st != GS_OK) return st;
  }
  // The id list above lives in host memory owned by this frame.
  return from_hip(hipStreamSynchronize(h->stream), "segmented sort");
}
