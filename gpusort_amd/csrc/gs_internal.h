// Host-side internal interface between the C ABI (capi.hip) and the kernel
// translation units.  Nothing here crosses the library boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace gs {

// Key kinds of the reference's instantiation matrix (src/cuda_utils.cu:257-262).
enum KeyKind : int { kI32 = 0, kU32 = 1, kF32 = 2, kI64 = 3, kU64 = 4, kF64 = 5 };

inline int key_bytes(int kind) { return kind <= kF32 ? 4 : 8; }

// Transform parameters for an unsigned image of width 4 or 8 bytes; the
// kernels take the 32- or 64-bit half that matches the key width.
struct XformParams {
  uint64_t float_mask;
  uint64_t sign_flip;
  uint64_t desc_mask;
};

XformParams make_xform(int kind, bool descending);

constexpr int kMaxPasses = 8;
constexpr uint32_t kSkipPass = 0xFFFFFFFFu;

// Device-resident control block of one radix sort: which ping-pong buffer
// each pass reads (decided on the device from the digit histograms, so no
// host round trip is needed), the tile tickets of each pass, and the buffer
// that holds the result.
struct SortCtrl {
  uint32_t src[kMaxPasses];
  uint32_t tile_counter[kMaxPasses];
  uint32_t final_sel;
  uint32_t passes_run;
  // Tile tickets per workgroup class (blockIdx % 8) when a pass hands out
  // tiles in XCD groups.
  uint32_t class_counter[kMaxPasses][8];
  uint32_t pad[14];
};

// Optional per-launch timing: when a launcher is given a timer it brackets
// each kernel with events recorded on the launch stream.  `bytes` is the
// algorithmic HBM traffic of the launch (inputs read + outputs written once).
class LaunchTimer {
 public:
  virtual ~LaunchTimer() {}
  virtual void begin(const char* name, uint64_t bytes, hipStream_t stream) = 0;
  virtual void end(hipStream_t stream) = 0;
};

// Tile hand-out of the onesweep passes.  Grouped hand-out (consecutive tiles
// on one XCD) needs about 8*G resident workgroups of the pass for its
// look-back to progress, which only a pass that has the GPU to itself is sure
// of; the arrival-order ticket progresses whatever else runs.  kTilesByStream:
// grouped on the null stream, arrival order on any other stream.
enum TilePolicy : int { kTilesByStream = -1, kTilesArrival = 0, kTilesGrouped = 1 };

// ---------------------------------------------------------------- radix sort
// Onesweep LSD radix sort of n keys (key_bytes 4/8) with an optional payload
// (val_bytes 0/4/8), stable.  keys/vals are sorted in place; the workspace
// holds the ping-pong buffers, histograms and look-back status words.
size_t radix_workspace_bytes(int key_bytes, int val_bytes, uint64_t n);
hipError_t radix_sort(int key_bytes, int val_bytes, void* keys, void* vals, uint64_t n, const XformParams& xf,
                      void* workspace, size_t workspace_bytes, hipStream_t stream, LaunchTimer* timer = nullptr,
                      int tile_policy = kTilesByStream);

// Segmented onesweep: many segments of one array sorted by one launch per
// digit.  Segment slot s is [seg_range[2s], seg_range[2s+1]) of the array,
// shorter than 2^30 keys, cut into tiles of radix_seg_tile_keys() keys numbered
// segment after segment (seg_first[s] = its first tile, tile_seg[t] = the
// slot of tile t); slots [0, n_multi) are the ones with more than one tile.
// All maps are device arrays.  n_total bounds every element offset (the
// ping-pong buffer in the workspace mirrors the whole array); n_keys is the
// number of keys in the segments.
struct SegTiles {
  int geo;  // tile geometry (radix_seg_tile_keys): 0 = the large one, 1-3 = the small-tile class's
  const uint32_t* tile_seg;
  const uint32_t* seg_first;
  const uint64_t* seg_range;
  uint64_t tiles;
  uint32_t nslots;
  uint32_t n_multi;
  uint64_t n_total;
  uint64_t n_keys;
};
// Stable multisplit of n keys (and payloads) into nspl + 1 <= 256 buckets by
// nspl sorted splitters (device memory, key type): bucket = number of
// splitters < key (upper = 1) or <= key (upper = 0), in the transform's
// order.  Writes out_keys/out_vals bucket by bucket, input order within a
// bucket, and the 256 exclusive bucket starts to d_bucket_base (device).
size_t multisplit_workspace_bytes(int key_bytes, int val_bytes, uint64_t n);
hipError_t multisplit(int key_bytes, int val_bytes, const void* keys, const void* vals, uint64_t n,
                      const void* d_splitters, uint32_t nspl, int upper, const XformParams& xf, void* out_keys,
                      void* out_vals, uint64_t* d_bucket_base, void* workspace, size_t workspace_bytes,
                      hipStream_t stream, LaunchTimer* timer = nullptr, int tile_policy = kTilesByStream);

// Keys per tile of segmented-sort geometry `geo` (0 or 1).
uint32_t radix_seg_tile_keys(int key_bytes, int val_bytes, int geo);

// Size classes of the batched segmented sort: segments of at most small_max
// keys (<= block_sort_max) are sorted by the LDS bitonic kernel, those of at
// most medium_max by the segmented onesweep with the small tiles, the rest
// with the large tiles.
struct SegConfig {
  uint32_t small_max;
  uint64_t medium_max;
  uint32_t medium_geometry = 0;  // small-tile class: 0 default, 1 half keys per thread, 2 twice the threads
};
SegConfig default_seg_config(int key_bytes, int val_bytes);

size_t radix_segments_workspace_bytes(int key_bytes, int val_bytes, uint64_t n_total, uint64_t tiles,
                                      uint32_t nslots);
hipError_t radix_sort_segments(int key_bytes, int val_bytes, void* keys, void* vals, const SegTiles& tiles,
                               const XformParams& xf, void* workspace, size_t workspace_bytes, hipStream_t stream,
                               LaunchTimer* timer = nullptr, int tile_policy = kTilesByStream);

// ---------------------------------------------------------------- block sort
// Single-workgroup stable sort of n <= block_sort_max(key_bytes, val_bytes)
// keys held in LDS (bitonic network over (key, original index) pairs).
uint32_t block_sort_max(int key_bytes, int val_bytes);
hipError_t block_sort(int key_bytes, int val_bytes, void* keys, void* vals, uint32_t n, const XformParams& xf,
                      hipStream_t stream);

// Batched variant: workgroup b sorts segment s = d_seg_ids[b] (or b when
// d_seg_ids is null), i.e. keys[off[s] .. off[s+1]), which must be no longer
// than block_sort_max.  max_len (0: unknown) bounds the segments' length;
// up to kBlockSortSmall keys a smaller workgroup with a smaller LDS tile is
// used, so more of them share a CU.
constexpr uint32_t kBlockSortSmall = 1024;
hipError_t block_sort_segments(int key_bytes, int val_bytes, void* keys, void* vals, const uint64_t* d_offsets,
                               const uint32_t* d_seg_ids, uint32_t nblocks, const XformParams& xf,
                               hipStream_t stream, uint32_t max_len = 0);

// ---------------------------------------------------------------- merge path
// One 2-way merge job: out[out_off ..) = merge(a[a_off, a_off+a_len), b[b_off, b_off+b_len)),
// ties taken from a first (std::merge semantics).  Offsets index the source
// and destination buffers given to merge_jobs.
struct MergeJob {
  uint64_t a_off, a_len, b_off, b_len, out_off;
};

// Device record of a job: the host record plus the index of its first tile.
struct MergeJobDev {
  uint64_t a_off, a_len, b_off, b_len, out_off;
  uint32_t tile0;
  uint32_t pad;
};

// Output elements per merge tile.
uint32_t merge_tile_elems();
// Fills out_jobs[0..njobs) and the tile -> job table; returns the tile count.
uint32_t merge_plan(const MergeJob* jobs, uint32_t njobs, MergeJobDev* out_jobs, std::vector<uint32_t>* tile_job);
// Runs every job: run a of a job is read from keys_a/vals_a + a_off, run b
// from keys_b/vals_b + b_off.  d_split needs ntiles entries of scratch.
// lex_vxf: when given (and there is a payload) elements compare as (key,
// payload) pairs, the payload ordered by this transform.
hipError_t merge_tiles(int key_bytes, int val_bytes, const void* keys_a, const void* keys_b, const void* vals_a,
                       const void* vals_b, void* dst_keys, void* dst_vals, const MergeJobDev* d_jobs,
                       const uint32_t* d_tile_job, uint32_t ntiles, const XformParams& xf, uint64_t* d_split,
                       hipStream_t stream, const XformParams* lex_vxf = nullptr);

// One merge-split of SampleSort's small-input bitonic network: rank list a
// (own) and b (partner), both sorted, in one device buffer.
struct NetCut {
  uint64_t a_off, a_len, b_off, b_len;
  uint32_t keep_low, pad;
};
// Kept length, and the merge-path split (a_split + b_split = the kept
// prefix's length, or the length of what precedes the kept suffix).
struct NetCutResult {
  uint64_t keep, a_split, b_split;
  uint32_t prefix, pad;
};
hipError_t net_cuts(int key_bytes, int val_bytes, const void* keys, const void* vals, const NetCut* d_ops,
                    uint32_t nops, const XformParams& xf, const XformParams* lex_vxf, int increasing,
                    NetCutResult* d_out, hipStream_t stream);

// ---------------------------------------------------------------- k-way merge
// One destination of a single-pass k-way merge: m <= kKwayMax sorted runs
// (offsets into the source buffer, in list order: equal elements come from
// the earlier run first) merged into [out_off, out_off + sum of lengths) of
// the destination buffer.  samp_base .. nparts are filled by kway_plan.
constexpr uint32_t kKwayMax = 8;
struct KwayDest {
  uint64_t run_off[kKwayMax];
  uint64_t run_len[kKwayMax];
  uint64_t out_off;
  uint32_t m;
  uint32_t samp_base;  // first sample of this destination
  uint32_t bnd_base;   // first of its nparts + 1 partition boundaries
  uint32_t part_base;  // first of its partitions
  uint32_t nparts;
  uint32_t pad;
};
// Sizes of a k-way pass: destinations, samples, partition boundaries and
// partitions.  The device copy (KwayTotals) holds the exact counts; the host
// sizes launches and scratch by the same fields (exact, or upper bounds when
// the plan is built on the device).
struct KwaySizes {
  uint32_t ndests = 0, nsamples = 0, nbounds = 0, nparts = 0;
};
using KwayTotals = KwaySizes;
// Host plan of a k-way pass: kway_plan fills each destination's sample,
// boundary and partition bases and the totals.
struct KwayPlan {
  std::vector<KwayDest> dests;
  KwaySizes totals;
};
void kway_plan(KwayPlan* plan);
// Upper bounds of a pass over n_elems elements in ndests destinations.
KwaySizes kway_upper_bounds(uint64_t n_elems, uint32_t ndests);
// Output elements per sample and samples per partition of the pass.
constexpr uint32_t kKwaySampleStride = 128;
constexpr uint32_t kKwaySamplesPerPart = 16;
// Device scratch of a pass (partition boundaries and compact samples).
size_t kway_scratch_bytes(const KwaySizes& sizes, int kb, int vb);
// Device tables of a pass: the destinations (with their bases), the exact
// totals, and the scratch.
struct KwayDevice {
  const KwayDest* dests;
  const KwayTotals* totals;
  void* scratch;
};
// Merges every destination from (sk, sv) into (dk, dv); `sizes` bounds the
// launches.  lex_vxf: when given (and there is a payload) elements compare as
// (key, payload) pairs.
// With a timer: one record for the boundary kernels ("kway_setup") and one
// for the merge kernel ("kway_merge_kernel", `elems` elements read and
// written).
hipError_t kway_merge(int kb, int vb, const void* sk, const void* sv, void* dk, void* dv, const KwaySizes& sizes,
                      const KwayDevice& dev, const XformParams& xf, const XformParams* lex_vxf, hipStream_t stream,
                      LaunchTimer* timer = nullptr, uint64_t elems = 0);

// ---------------------------------------------------------------- HykSort rounds
// HykSort's rounds over p virtual ranks with every step on the device
// (hyk_device.hip).  The runs [off[r], off[r+1]) of kbuf[cur] / vbuf[cur]
// must be sorted (lexicographically by (key, payload) when lex); on return
// cur names the buffer holding the result.
struct HykRun {
  int kb, vb;
  bool lex;
  XformParams kx, vx;
  char* kbuf[2];
  char* vbuf[2];
  int cur;
  uint32_t p, k_way, s;
  uint64_t n;
  // Pinned host memory of at least hyk_host_bytes(): the selection summaries
  // and the report come back through it (one copy each).
  void* host;
};
// What the rounds leave for the report: per round, the bucket bounds
// (block of p * bnd_stride words, rank r at r * (kway + 2): run start,
// lower_bound of each splitter, run end, element offsets) and the splitters
// (block of p * spl_stride raw keys / payloads, group g's at g * kway), the
// selection iterations, and the final rank offsets.
struct HykResult {
  uint32_t rounds = 0;
  std::vector<uint32_t> iterations, kway, n_p;
  std::vector<uint64_t> bounds, final_off;
  std::vector<char> spl_k, spl_v;
  size_t bnd_stride = 0, spl_stride = 0;
};
bool hyk_device_supported(uint32_t p, uint32_t k_way, uint32_t s);
size_t hyk_device_workspace_bytes(uint32_t p, uint32_t k_way, uint32_t s, uint64_t n, int kb, int vb);
size_t hyk_host_bytes(uint32_t p, uint32_t k_way, uint32_t s, int kb, int vb);
hipError_t hyk_rounds_device(HykRun* run, const uint64_t* h_off, void* workspace, size_t workspace_bytes,
                             hipStream_t stream, LaunchTimer* timer, HykResult* res);

// ---------------------------------------------------------------- splitting
// Bucket boundaries of sorted runs.  Run r is keys[run_off[r] .. run_off[r+1])
// and is cut by splitters[spl_first[r] .. spl_first[r] + nspl).  Output:
// bounds[r*(nspl+2) + j], j = 0..nspl+1: run start, then run start +
// lower_bound (upper_bound if `upper`) of each splitter, then run end.
// With lex_vxf (and a payload of val_bytes), runs and splitters are (key,
// payload) pairs compared lexicographically, the payload ordered by lex_vxf.
hipError_t bucket_bounds(int key_bytes, const void* keys, const uint64_t* d_run_off, uint32_t nruns,
                         const void* d_splitters, const uint32_t* d_spl_first, uint32_t nspl, int upper,
                         const XformParams& xf, uint64_t* d_bounds, hipStream_t stream, int val_bytes = 0,
                         const void* vals = nullptr, const void* d_spl_vals = nullptr,
                         const XformParams* lex_vxf = nullptr);
// out[i] = src[idx[i]] for n elements of 4 or 8 bytes.
hipError_t gather_elems(int bytes, const void* src, const uint64_t* d_idx, void* out, uint32_t n, hipStream_t stream);
// Sets *d_flag (not cleared here) when a[0..n) is not non-decreasing under xf.
hipError_t unsorted_flag(int bytes, const void* a, uint64_t n, const XformParams& xf, uint32_t* d_flag,
                         hipStream_t stream);

// ---------------------------------------------------------------- utilities
hipError_t fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream);
hipError_t copy_bytes(void* dst, const void* src, uint64_t bytes, hipStream_t stream);
hipError_t gather_values(int val_bytes, const void* src, const uint64_t* idx, void* dst, uint64_t n,
                         hipStream_t stream);
hipError_t pack_pairs(int key_bytes, const void* keys, int val_bytes, const void* vals, void* out, uint64_t n,
                      uint32_t stride_bytes, uint32_t val_offset, hipStream_t stream);
hipError_t iota_u64(uint64_t* out, uint64_t n, uint64_t base, hipStream_t stream);

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace gs
