// Batched segmented sort (gs_sort_segments).
//
// The reference sorts one array per call (CudaUtils::Sort / SortByKey,
// src/cuda_utils.cu:97-243); a caller with many arrays pays one launch
// sequence and one synchronisation per array.  Here every segment of one
// call is sorted by a fixed number of launches, whatever the count:
//
//   small  (<= one LDS tile, block_sort_max): one workgroup per segment,
//          bitonic network in LDS, all small segments in one launch
//   radix  (up to 2^30 keys): segmented onesweep -- one histogram launch,
//          one scan launch and one launch per digit for all of them; tiles
//          never straddle segments and the look-back stops at the segment's
//          first tile.  Multi-tile segments are numbered first.
//   huge   (>= 2^30 keys): the single-array radix sort, one per segment
//
// Every class is stable with a payload, so each segment's result equals a
// separate gs_sort_pairs call on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/gpusort_hip.h"
#include "gs_capi_impl.h"
#include "gs_internal.h"

using namespace gs;

namespace gs {

constexpr uint64_t kSegRadixMax = 1ull << 30;

// Sorts segment s = [off[s], off[s+1]) of keys/vals (host offsets) with no
// host synchronisation: the launches are queued on h->stream.  The device
// maps are uploaded from `meta`, which the caller keeps alive until the
// stream has passed this point.  Segments of 2^30 keys or more, which the
// segmented radix cannot take, are left to the caller (returned in *huge).
gs_status sort_segments_async(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals,
                              const uint64_t* off, uint32_t nseg, int descending, std::vector<unsigned char>* meta,
                              std::vector<uint32_t>* huge) {
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  const XformParams xf = make_xform((int)kdt, descending != 0);
  hipError_t e = hipSuccess;

  // Size classes.
  const uint32_t small_max = block_sort_max(kb, vb);
  const uint64_t tile = radix_tile_keys(kb, vb);
  std::vector<uint32_t> small_ids, multi_ids, single_ids;
  huge->clear();
  for (uint32_t s = 0; s < nseg; ++s) {
    const uint64_t len = off[s + 1] - off[s];
    if (len <= 1) continue;
    if (len <= small_max)
      small_ids.push_back(s);
    else if (len >= kSegRadixMax)
      huge->push_back(s);
    else if (len > tile)
      multi_ids.push_back(s);
    else
      single_ids.push_back(s);
  }

  // Device maps of the radix class (multi-tile slots first), the small ids
  // and the offsets, uploaded in one copy.
  const uint32_t nslots = (uint32_t)(multi_ids.size() + single_ids.size());
  uint64_t tiles = 0, n_keys = 0;
  for (uint32_t s : multi_ids) tiles += (off[s + 1] - off[s] + tile - 1) / tile;
  tiles += single_ids.size();
  size_t o = 0;
  const size_t off_tile_seg = o;
  o = align_up(o + tiles * 4, 256);
  const size_t off_seg_first = o;
  o = align_up(o + (size_t)nslots * 4, 256);
  const size_t off_seg_range = o;
  o = align_up(o + (size_t)nslots * 16, 256);
  const size_t off_small = o;
  o = align_up(o + small_ids.size() * 4, 256);
  const size_t off_offsets = o;
  o = align_up(o + (size_t)(nseg + 1) * 8, 256);
  meta->assign(o, 0);
  {
    unsigned char* b = meta->data();
    uint32_t* tile_seg = reinterpret_cast<uint32_t*>(b + off_tile_seg);
    uint32_t* seg_first = reinterpret_cast<uint32_t*>(b + off_seg_first);
    uint64_t* seg_range = reinterpret_cast<uint64_t*>(b + off_seg_range);
    uint64_t t = 0;
    uint32_t slot = 0;
    for (const std::vector<uint32_t>* ids : {&multi_ids, &single_ids}) {
      for (uint32_t s : *ids) {
        const uint64_t len = off[s + 1] - off[s];
        const uint64_t nt = (len + tile - 1) / tile;
        seg_first[slot] = (uint32_t)t;
        seg_range[2 * slot] = off[s];
        seg_range[2 * slot + 1] = off[s + 1];
        for (uint64_t j = 0; j < nt; ++j) tile_seg[t++] = slot;
        n_keys += len;
        ++slot;
      }
    }
    if (!small_ids.empty()) memcpy(b + off_small, small_ids.data(), small_ids.size() * 4);
    memcpy(b + off_offsets, off, (size_t)(nseg + 1) * 8);
  }
  if (small_ids.empty() && nslots == 0) return GS_OK;
  void* dmeta = nullptr;
  e = h->arena.get(Arena::kSegAux, meta->size(), h->stream, &dmeta);
  if (e != hipSuccess) return from_hip(e, "segment maps");
  e = hipMemcpyAsync(dmeta, meta->data(), meta->size(), hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) return from_hip(e, "upload segment maps");
  unsigned char* dm = reinterpret_cast<unsigned char*>(dmeta);

  if (!small_ids.empty()) {
    e = block_sort_segments(kb, vb, d_keys, d_vals, reinterpret_cast<const uint64_t*>(dm + off_offsets),
                            reinterpret_cast<const uint32_t*>(dm + off_small), (uint32_t)small_ids.size(), xf,
                            h->stream);
    if (e != hipSuccess) return from_hip(e, "small segments");
  }
  if (nslots) {
    SegTiles T;
    T.tile_seg = reinterpret_cast<const uint32_t*>(dm + off_tile_seg);
    T.seg_first = reinterpret_cast<const uint32_t*>(dm + off_seg_first);
    T.seg_range = reinterpret_cast<const uint64_t*>(dm + off_seg_range);
    T.tiles = tiles;
    T.nslots = nslots;
    T.n_multi = (uint32_t)multi_ids.size();
    T.n_total = off[nseg];
    T.n_keys = n_keys;
    const size_t ws_bytes = radix_segments_workspace_bytes(kb, vb, T.n_total, tiles, nslots);
    void* ws = nullptr;
    e = h->arena.get(Arena::kSegWork, ws_bytes, h->stream, &ws);
    if (e != hipSuccess) return from_hip(e, "segmented radix workspace");
    e = radix_sort_segments(kb, vb, d_keys, d_vals, T, xf, ws, ws_bytes, h->stream, h->timer());
    if (e != hipSuccess) return from_hip(e, "segmented radix sort");
  }
  return GS_OK;
}

}  // namespace gs

extern "C" gs_status gs_sort_segments(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals,
                                      const uint64_t* d_offsets, uint32_t nseg, int descending) {
  if (!h) return fail(GS_INVALID_ARG, "null handle");
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (vdt != GS_NONE && !dtype_ok(vdt)) return fail(GS_INVALID_ARG, "unsupported value type %d", (int)vdt);
  if (nseg == 0) return GS_OK;
  if (!d_offsets || !d_keys || (vdt != GS_NONE && !d_vals)) return fail(GS_INVALID_ARG, "null pointer");
  DeviceScope ds(h->device);
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  std::vector<uint64_t> off(nseg + 1);
  hipError_t e = hipMemcpyAsync(off.data(), d_offsets, (nseg + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return from_hip(e, "read segment offsets");
  for (uint32_t s = 0; s < nseg; ++s)
    if (off[s + 1] < off[s]) return fail(GS_INVALID_ARG, "segment offsets decrease at segment %u", s);
  std::vector<unsigned char> meta;
  std::vector<uint32_t> huge;
  gs_status st = sort_segments_async(h, kdt, d_keys, vdt, d_vals, off.data(), nseg, descending, &meta, &huge);
  if (st != GS_OK) return st;
  char* kbase = reinterpret_cast<char*>(d_keys);
  char* vbase = reinterpret_cast<char*>(d_vals);
  for (uint32_t s : huge) {
    const uint64_t len = off[s + 1] - off[s];
    st = sort_device(h, kdt, kbase + off[s] * kb, vdt, vb ? vbase + off[s] * vb : nullptr, len, descending);
    if (st != GS_OK) return st;
  }
  // The maps were copied from host memory owned by this frame.
  return from_hip(hipStreamSynchronize(h->stream), "segmented sort");
}
