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
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/gpusort_hip.h"
#include "gs_capi_impl.h"
#include "gs_internal.h"

using namespace gs;

namespace gs {

constexpr uint64_t kSegRadixMax = 1ull << 30;

SegConfig default_seg_config(int kb, int vb) { return SegConfig{block_sort_max(kb, vb), 0}; }

// Host side of one batched sort: the size classes of the segments and the
// device maps (uploaded from `meta`), independent of the key/payload buffers.
struct SegPlanHost {
  int kb = 4, vb = 0;
  uint32_t nseg = 0;
  uint64_t n_total = 0;
  std::vector<uint32_t> small_ids, multi_ids[2], single_ids[2], huge;
  uint32_t small_len = 0;  // longest segment of the LDS class
  struct Group {
    int geo = 0;  // tile geometry (radix_seg_tile_keys)
    uint32_t nslots = 0;
    uint64_t tiles = 0, n_keys = 0, tile_keys = 0;
    size_t off_tile_seg = 0, off_seg_first = 0, off_seg_range = 0;
  } grp[2];
  size_t off_small = 0, off_offsets = 0;
  std::vector<unsigned char> meta;
  size_t ws_bytes = 0;  // workspace of the radix groups
};

// Classifies the segments (radix groups g: 1 = small tiles, of the
// configured medium geometry, up to medium_max keys; 0 = large tiles;
// multi-tile segments first) and lays out the maps.
void build_seg_plan(int kb, int vb, const uint64_t* off, uint32_t nseg, const SegConfig& cfg, SegPlanHost* P) {
  P->kb = kb;
  P->vb = vb;
  P->nseg = nseg;
  P->n_total = off[nseg];
  const uint32_t small_max = std::min<uint32_t>(cfg.small_max, block_sort_max(kb, vb));
  P->grp[0].geo = 0;
  P->grp[1].geo = 1 + (int)std::min<uint32_t>(cfg.medium_geometry, 2u);
  for (uint32_t s = 0; s < nseg; ++s) {
    const uint64_t len = off[s + 1] - off[s];
    if (len <= 1) continue;
    if (len <= small_max) {
      P->small_ids.push_back(s);
      P->small_len = std::max(P->small_len, (uint32_t)len);
    } else if (len >= kSegRadixMax) {
      P->huge.push_back(s);
    } else {
      const int g = len <= cfg.medium_max ? 1 : 0;
      (len > radix_seg_tile_keys(kb, vb, P->grp[g].geo) ? P->multi_ids[g] : P->single_ids[g]).push_back(s);
    }
  }
  size_t o = 0;
  for (int g = 0; g < 2; ++g) {
    SegPlanHost::Group& G = P->grp[g];
    G.tile_keys = radix_seg_tile_keys(kb, vb, G.geo);
    G.nslots = (uint32_t)(P->multi_ids[g].size() + P->single_ids[g].size());
    for (uint32_t s : P->multi_ids[g]) G.tiles += (off[s + 1] - off[s] + G.tile_keys - 1) / G.tile_keys;
    G.tiles += P->single_ids[g].size();
    G.off_tile_seg = o;
    o = align_up(o + G.tiles * 4, 256);
    G.off_seg_first = o;
    o = align_up(o + (size_t)G.nslots * 4, 256);
    G.off_seg_range = o;
    o = align_up(o + (size_t)G.nslots * 16, 256);
    if (G.nslots)
      P->ws_bytes = std::max(P->ws_bytes, radix_segments_workspace_bytes(kb, vb, P->n_total, G.tiles, G.nslots));
  }
  P->off_small = o;
  o = align_up(o + P->small_ids.size() * 4, 256);
  P->off_offsets = o;
  o = align_up(o + (size_t)(nseg + 1) * 8, 256);
  P->meta.assign(o, 0);
  unsigned char* b = P->meta.data();
  for (int g = 0; g < 2; ++g) {
    SegPlanHost::Group& G = P->grp[g];
    uint32_t* tile_seg = reinterpret_cast<uint32_t*>(b + G.off_tile_seg);
    uint32_t* seg_first = reinterpret_cast<uint32_t*>(b + G.off_seg_first);
    uint64_t* seg_range = reinterpret_cast<uint64_t*>(b + G.off_seg_range);
    uint64_t t = 0;
    uint32_t slot = 0;
    for (const std::vector<uint32_t>* ids : {&P->multi_ids[g], &P->single_ids[g]}) {
      for (uint32_t s : *ids) {
        const uint64_t len = off[s + 1] - off[s];
        const uint64_t nt = (len + G.tile_keys - 1) / G.tile_keys;
        seg_first[slot] = (uint32_t)t;
        seg_range[2 * slot] = off[s];
        seg_range[2 * slot + 1] = off[s + 1];
        for (uint64_t j = 0; j < nt; ++j) tile_seg[t++] = slot;
        G.n_keys += len;
        ++slot;
      }
    }
  }
  if (!P->small_ids.empty()) memcpy(b + P->off_small, P->small_ids.data(), P->small_ids.size() * 4);
  memcpy(b + P->off_offsets, off, (size_t)(nseg + 1) * 8);
}

bool seg_plan_empty(const SegPlanHost& P) {
  return P.small_ids.empty() && P.grp[0].nslots == 0 && P.grp[1].nslots == 0;
}

// Kernel launches only (no copies from host memory, no allocation, no
// synchronisation): dm is the device copy of P.meta, ws holds P.ws_bytes.
hipError_t launch_seg_plan(const SegPlanHost& P, const unsigned char* dm, void* ws, void* d_keys, void* d_vals,
                           const XformParams& xf, hipStream_t stream, LaunchTimer* timer,
                           int tile_policy = kTilesByStream) {
  hipError_t e = hipSuccess;
  if (!P.small_ids.empty()) {
    e = block_sort_segments(P.kb, P.vb, d_keys, d_vals, reinterpret_cast<const uint64_t*>(dm + P.off_offsets),
                            reinterpret_cast<const uint32_t*>(dm + P.off_small), (uint32_t)P.small_ids.size(), xf,
                            stream, P.small_len);
    if (e != hipSuccess) return e;
  }
  for (int g = 1; g >= 0; --g) {
    const SegPlanHost::Group& G = P.grp[g];
    if (!G.nslots) continue;
    SegTiles T;
    T.geo = G.geo;
    T.tile_seg = reinterpret_cast<const uint32_t*>(dm + G.off_tile_seg);
    T.seg_first = reinterpret_cast<const uint32_t*>(dm + G.off_seg_first);
    T.seg_range = reinterpret_cast<const uint64_t*>(dm + G.off_seg_range);
    T.tiles = G.tiles;
    T.nslots = G.nslots;
    T.n_multi = (uint32_t)P.multi_ids[g].size();
    T.n_total = P.n_total;
    T.n_keys = G.n_keys;
    e = radix_sort_segments(P.kb, P.vb, d_keys, d_vals, T, xf, ws, P.ws_bytes, stream, timer, tile_policy);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

gs_status sort_segments_async(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals,
                              const uint64_t* off, uint32_t nseg, int descending, std::vector<unsigned char>* meta,
                              std::vector<uint32_t>* huge, const SegConfig* config) {
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  const SegConfig cfg = config ? *config : h->seg_config(kb, vb);
  SegPlanHost P;
  build_seg_plan(kb, vb, off, nseg, cfg, &P);
  *huge = P.huge;
  meta->swap(P.meta);
  if (seg_plan_empty(P)) return GS_OK;
  void *dmeta = nullptr, *ws = nullptr;
  hipError_t e = h->arena.get(Arena::kSegAux, meta->size(), h->stream, &dmeta);
  if (e == hipSuccess) e = hipMemcpyAsync(dmeta, meta->data(), meta->size(), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess && P.ws_bytes) e = h->arena.get(Arena::kSegWork, P.ws_bytes, h->stream, &ws);
  if (e != hipSuccess) return from_hip(e, "segment maps and workspace");
  e = launch_seg_plan(P, reinterpret_cast<unsigned char*>(dmeta), ws, d_keys, d_vals,
                      make_xform((int)kdt, descending != 0), h->stream, h->timer(), h->tile_policy);
  return from_hip(e, "segmented sort");
}

}  // namespace gs

extern "C" gs_status gs_sort_segments(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals,
                                      const uint64_t* d_offsets, uint32_t nseg, int descending) {
  if (!h) return fail(GS_INVALID_ARG, "null handle");
  if (!dtype_ok(kdt)) return fail(GS_INVALID_ARG, "unsupported key type %d", (int)kdt);
  if (vdt != GS_NONE && !dtype_ok(vdt)) return fail(GS_INVALID_ARG, "unsupported value type %d", (int)vdt);
  if (nseg == 0) return GS_OK;
  if (!d_offsets) return fail(GS_INVALID_ARG, "null pointer");
  DeviceScope ds(h->device);
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  std::vector<uint64_t> off(nseg + 1);
  hipError_t e = hipMemcpyAsync(off.data(), d_offsets, (nseg + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return from_hip(e, "read segment offsets");
  for (uint32_t s = 0; s < nseg; ++s)
    if (off[s + 1] < off[s]) return fail(GS_INVALID_ARG, "segment offsets decrease at segment %u", s);
  if (off[nseg] == off[0]) return GS_OK;  // every segment empty: nothing is read (null pointers allowed)
  if (!d_keys || (vdt != GS_NONE && !d_vals)) return fail(GS_INVALID_ARG, "null pointer");
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

// ------------------------------------------------------------ plans and graphs
//
// A plan fixes the segment offsets, types and direction of a batched sort:
// the classification, device maps and workspace are made once, so running it
// is kernel launches only and can be captured into a HIP graph and replayed.

struct gs_seg_plan {
  gs_handle* h = nullptr;
  gs_dtype kdt = GS_U32, vdt = GS_NONE;
  XformParams xf{};
  gs::SegPlanHost host;
  void* d_meta = nullptr;
  void* d_ws = nullptr;
  hipStream_t own_stream = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  void* graph_keys = nullptr;
  void* graph_vals = nullptr;
  hipStream_t stream() const { return h->stream ? h->stream : own_stream; }
  // The plan's own stream stands in for the handle's null stream (which a
  // graph cannot capture), so it keeps the null stream's grouped hand-out.
  int tile_policy() const {
    return (h->tile_policy == kTilesByStream && h->stream == nullptr) ? (int)kTilesGrouped : h->tile_policy;
  }
  void drop_graph() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

extern "C" {

gs_status gs_get_seg_config(gs_handle* h, gs_dtype kdt, gs_dtype vdt, gs_seg_config* out) {
  if (!h || !out) return fail(GS_INVALID_ARG, "null argument");
  if (!dtype_ok(kdt) || (vdt != GS_NONE && !dtype_ok(vdt))) return fail(GS_INVALID_ARG, "unsupported type");
  const SegConfig c = h->seg_config(dtype_bytes(kdt), val_bytes_of(vdt));
  out->small_max = c.small_max;
  out->medium_max = c.medium_max;
  out->medium_geometry = c.medium_geometry;
  return GS_OK;
}

gs_status gs_set_seg_config(gs_handle* h, gs_dtype kdt, gs_dtype vdt, const gs_seg_config* cfg) {
  if (!h || !cfg) return fail(GS_INVALID_ARG, "null argument");
  if (!dtype_ok(kdt) || (vdt != GS_NONE && !dtype_ok(vdt))) return fail(GS_INVALID_ARG, "unsupported type");
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  if (cfg->small_max > block_sort_max(kb, vb))
    return fail(GS_INVALID_ARG, "small_max %u exceeds the LDS tile of %u keys", cfg->small_max, block_sort_max(kb, vb));
  if (cfg->medium_geometry > 2) return fail(GS_INVALID_ARG, "medium_geometry %u is not 0, 1 or 2", cfg->medium_geometry);
  h->seg_tuned[kb * 16 + vb] = SegConfig{cfg->small_max, cfg->medium_max, cfg->medium_geometry};
  return GS_OK;
}

gs_status gs_seg_plan_create(gs_handle* h, gs_dtype kdt, gs_dtype vdt, const uint64_t* h_offsets, uint32_t nseg,
                             int descending, gs_seg_plan** out) {
  if (!h || !out) return fail(GS_INVALID_ARG, "null argument");
  *out = nullptr;
  if (!dtype_ok(kdt) || (vdt != GS_NONE && !dtype_ok(vdt))) return fail(GS_INVALID_ARG, "unsupported type");
  if (nseg && !h_offsets) return fail(GS_INVALID_ARG, "null offsets");
  for (uint32_t s = 0; s < nseg; ++s)
    if (h_offsets[s + 1] < h_offsets[s]) return fail(GS_INVALID_ARG, "segment offsets decrease at segment %u", s);
  DeviceScope ds(h->device);
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);
  gs_seg_plan* p = new gs_seg_plan();
  p->h = h;
  p->kdt = kdt;
  p->vdt = vdt;
  p->xf = make_xform((int)kdt, descending != 0);
  static const uint64_t zero = 0;
  build_seg_plan(kb, vb, nseg ? h_offsets : &zero, nseg, h->seg_config(kb, vb), &p->host);
  if (!p->host.huge.empty()) {
    delete p;
    return fail(GS_UNSUPPORTED, "a plan takes segments shorter than 2^30 keys; use gs_sort_segments");
  }
  hipError_t e = hipStreamCreate(&p->own_stream);
  if (e == hipSuccess) e = hipMalloc(&p->d_meta, p->host.meta.size());
  if (e == hipSuccess && p->host.ws_bytes) e = hipMalloc(&p->d_ws, p->host.ws_bytes);
  if (e == hipSuccess) e = hipMemcpy(p->d_meta, p->host.meta.data(), p->host.meta.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    gs_seg_plan_destroy(p);
    return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  }
  *out = p;
  return GS_OK;
}

gs_status gs_seg_plan_run(gs_seg_plan* p, void* d_keys, void* d_vals) {
  if (!p) return fail(GS_INVALID_ARG, "null plan");
  if (p->host.n_total && (!d_keys || (p->vdt != GS_NONE && !d_vals))) return fail(GS_INVALID_ARG, "null pointer");
  DeviceScope ds(p->h->device);
  return from_hip(launch_seg_plan(p->host, reinterpret_cast<unsigned char*>(p->d_meta), p->d_ws, d_keys, d_vals, p->xf,
                                  p->stream(), p->h->timer(), p->tile_policy()),
                  "planned segmented sort");
}

gs_status gs_seg_plan_launch_graph(gs_seg_plan* p, void* d_keys, void* d_vals) {
  if (!p) return fail(GS_INVALID_ARG, "null plan");
  if (p->host.n_total && (!d_keys || (p->vdt != GS_NONE && !d_vals))) return fail(GS_INVALID_ARG, "null pointer");
  DeviceScope ds(p->h->device);
  hipStream_t st = p->stream();
  if (!p->exec || d_keys != p->graph_keys || d_vals != p->graph_vals) {
    p->drop_graph();
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return from_hip(e, "begin graph capture");
    const hipError_t le = launch_seg_plan(p->host, reinterpret_cast<unsigned char*>(p->d_meta), p->d_ws, d_keys,
                                          d_vals, p->xf, st, nullptr, p->tile_policy());
    e = hipStreamEndCapture(st, &p->graph);
    if (le != hipSuccess) return from_hip(le, "capture planned sort");
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      p->drop_graph();
      return from_hip(e, "instantiate graph");
    }
    p->graph_keys = d_keys;
    p->graph_vals = d_vals;
  }
  return from_hip(hipGraphLaunch(p->exec, st), "graph launch");
}

gs_status gs_seg_plan_destroy(gs_seg_plan* p) {
  if (!p) return GS_OK;
  DeviceScope ds(p->h->device);
  if (p->own_stream) hipStreamSynchronize(p->own_stream);
  if (p->h->stream) hipStreamSynchronize(p->h->stream);
  p->drop_graph();
  if (p->d_meta) hipFree(p->d_meta);
  if (p->d_ws) hipFree(p->d_ws);
  if (p->own_stream) hipStreamDestroy(p->own_stream);
  delete p;
  return GS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ autotuner

namespace {

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Uniform in (0, 1] from draw i of a splitmix64 stream.
double unit_draw(uint64_t seed, uint64_t i) {
  return (double)((mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull) >> 11) + 1) * (1.0 / 9007199254740992.0);
}

}  // namespace

extern "C" gs_status gs_autotune_segments(gs_handle* h, gs_dtype kdt, gs_dtype vdt, uint64_t total_keys,
                                          double median, double sigma, gs_seg_config* best, gs_tune_row* rows,
                                          uint32_t max_rows, uint32_t* n_rows) {
  if (!h) return fail(GS_INVALID_ARG, "null handle");
  if (!dtype_ok(kdt) || (vdt != GS_NONE && !dtype_ok(vdt))) return fail(GS_INVALID_ARG, "unsupported type");
  if (total_keys == 0 || total_keys > (1ull << 30) || !(median >= 1.0) || !(sigma >= 0.0))
    return fail(GS_INVALID_ARG, "need 0 < total_keys <= 2^30, median >= 1, sigma >= 0");
  DeviceScope ds(h->device);
  const int kb = dtype_bytes(kdt), vb = val_bytes_of(vdt);

  // Synthetic batch: lognormal lengths (Box-Muller on a counter stream).
  std::vector<uint64_t> off(1, 0);
  const double two_pi = 6.283185307179586;
  for (uint64_t i = 0; off.back() < total_keys; ++i) {
    const double z = std::sqrt(-2.0 * std::log(unit_draw(77, 2 * i))) * std::cos(two_pi * unit_draw(77, 2 * i + 1));
    double len = std::round(std::exp(std::log(median) + sigma * z));
    len = std::min(std::max(len, 1.0), (double)(1u << 22));
    off.push_back(off.back() + (uint64_t)len);
  }
  const uint64_t n = off.back();
  const uint32_t nseg = (uint32_t)off.size() - 1;
  void *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr;
  hipError_t e = hipMalloc(&k0, n * kb);
  if (e == hipSuccess) e = hipMalloc(&k1, n * kb);
  if (e == hipSuccess && vb) e = hipMalloc(&v0, n * vb);
  if (e == hipSuccess && vb) e = hipMalloc(&v1, n * vb);
  auto release = [&]() {
    hipFree(k0);
    hipFree(k1);
    if (v0) hipFree(v0);
    if (v1) hipFree(v1);
  };
  if (e != hipSuccess) {
    (void)hipGetLastError();
    release();
    return fail(GS_OUT_OF_MEMORY, "Couldn't allocate device vector");
  }
  gs_status st = gs_generate(h, kb == 4 ? 0 : 1, 5, k0, n, 0, 0.0, nullptr, 0);
  if (st == GS_OK && vb) st = gs_generate(h, vb == 4 ? 6 : 7, 0, v0, n, 0, 0.0, nullptr, 0);
  if (st != GS_OK) {
    release();
    return st;
  }

  const uint32_t cap = block_sort_max(kb, vb);
  const uint64_t t1 = radix_seg_tile_keys(kb, vb, 1);
  std::vector<SegConfig> cands;
  for (uint32_t sm : {cap, cap / 2, cap / 4})
    for (uint64_t md : {(uint64_t)0, 8 * t1, 64 * t1})
      for (uint32_t geo = 0; geo < (md ? 3u : 1u); ++geo) cands.push_back(SegConfig{sm, md, geo});
  const XformParams xf = make_xform((int)kdt, false);
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  std::vector<double> times(cands.size(), 0.0);
  for (size_t c = 0; c < cands.size() && e == hipSuccess; ++c) {
    SegPlanHost P;
    build_seg_plan(kb, vb, off.data(), nseg, cands[c], &P);
    void *dmeta = nullptr, *ws = nullptr;
    e = h->arena.get(Arena::kSegAux, P.meta.size(), h->stream, &dmeta);
    if (e == hipSuccess) e = hipMemcpyAsync(dmeta, P.meta.data(), P.meta.size(), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess && P.ws_bytes) e = h->arena.get(Arena::kSegWork, P.ws_bytes, h->stream, &ws);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    std::vector<double> ms;
    for (int r = 0; r < 4 && e == hipSuccess; ++r) {  // the first run is a warm-up
      e = hipMemcpyAsync(k1, k0, n * kb, hipMemcpyDeviceToDevice, h->stream);
      if (e == hipSuccess && vb) e = hipMemcpyAsync(v1, v0, n * vb, hipMemcpyDeviceToDevice, h->stream);
      if (e == hipSuccess) e = hipEventRecord(ev0, h->stream);
      if (e == hipSuccess)
        e = launch_seg_plan(P, reinterpret_cast<unsigned char*>(dmeta), ws, k1, v1, xf, h->stream, nullptr,
                            h->tile_policy);
      if (e == hipSuccess) e = hipEventRecord(ev1, h->stream);
      if (e == hipSuccess) e = hipEventSynchronize(ev1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, ev0, ev1);
      if (e == hipSuccess && r > 0) ms.push_back(t);
    }
    if (e == hipSuccess) {
      std::sort(ms.begin(), ms.end());
      times[c] = ms[ms.size() / 2];
    }
  }
  if (ev0) hipEventDestroy(ev0);
  if (ev1) hipEventDestroy(ev1);
  release();
  if (e != hipSuccess) return from_hip(e, "autotune");
  size_t bi = 0;
  for (size_t c = 1; c < cands.size(); ++c)
    if (times[c] < times[bi]) bi = c;
  h->seg_tuned[kb * 16 + vb] = cands[bi];
  if (best) {
    best->small_max = cands[bi].small_max;
    best->medium_max = cands[bi].medium_max;
    best->medium_geometry = cands[bi].medium_geometry;
  }
  if (n_rows) *n_rows = (uint32_t)std::min<size_t>(cands.size(), max_rows);
  for (size_t c = 0; rows && c < cands.size() && c < max_rows; ++c)
    rows[c] = gs_tune_row{gs_seg_config{cands[c].small_max, cands[c].medium_max, cands[c].medium_geometry}, times[c]};
  return GS_OK;
}
