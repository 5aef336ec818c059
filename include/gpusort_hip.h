/*
 * gpusort_hip.h -- C ABI of the MI355X (gfx950) per-rank sort path.
 *
 * This is the drop-in seam for the reference's GPU local sort and per-rank
 * stages.  Plain pointers and sizes only; every entry returns a gs_status and
 * never throws.  The C++ headers under include/gpusort/ wrap these entries in
 * the reference's template API (HykSort / SampleSort / HyperQuickSort /
 * CudaUtils / ParUtils) and re-throw errors as the reference's exceptions.
 *
 * Citations are files of monotone-RK/gpusort.
 */
#ifndef GPUSORT_HIP_H_
#define GPUSORT_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference reports errors as ThrustException /
 * MpiException messages (src/cuda_utils.cu:373-393, :415-441,
 * src/mpi_impl.cc:219-266); gs_last_error() returns the message. */
typedef enum gs_status {
  GS_OK = 0,
  GS_INVALID_ARG = 1,
  GS_OUT_OF_MEMORY = 2,
  GS_ADMISSION_REFUSED = 3,
  GS_DEVICE_ERROR = 4,
  GS_NO_DEVICE = 5,
  GS_UNSUPPORTED = 6,
  GS_EMPTY_INPUT = 7,
  GS_NOT_READY = 8 /* a wait timed out before the request finished */
} gs_status;

/* Key / value element types: the primitive types of the reference's
 * instantiation matrix (src/cuda_utils.cu:257-262, :288-318). */
typedef enum gs_dtype {
  GS_I32 = 0,
  GS_U32 = 1,
  GS_F32 = 2,
  GS_I64 = 3,
  GS_U64 = 4,
  GS_F64 = 5,
  GS_NONE = 255 /* no payload */
} gs_dtype;

/* Per-rank algorithms of the vector API. */
typedef enum gs_algorithm {
  GS_HYKSORT = 0,       /* include/helpers/hyksort_helper.h */
  GS_SAMPLESORT = 1,    /* include/helpers/samplesort_helper.h */
  GS_HYPERQUICKSORT = 2 /* include/helpers/hyperquicksort_helper.h */
} gs_algorithm;

typedef struct gs_handle gs_handle;

/* Library version string. */
const char* gs_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* gs_last_error(void);

/* Number of visible HIP devices.
 * Replaces CudaUtils::GetDeviceCount (include/cuda/cuda_utils.h:12,
 * src/cuda_utils.cu:81-88); IsHavingDevices is count > 0 (cuda_utils.cu:90-92). */
gs_status gs_get_device_count(int* count);

/* Creates a handle bound to `device` (its stream is the null stream until
 * gs_set_stream).  The handle owns the HBM workspace arena reused by every
 * call; device-pointer entries allocate only when the arena must grow.
 * Replaces SelectDevice (src/cuda_utils.cu:395-401). */
gs_status gs_create(gs_handle** out, int device);
gs_status gs_destroy(gs_handle* h);
gs_status gs_set_stream(gs_handle* h, void* hip_stream);

/* Tile hand-out of the handle's radix passes (sorts, segmented sorts,
 * multisplit).  A pass's tiles finish in tile order, so a tile may only wait
 * on tiles that started workgroups hold.  GS_TILES_AUTO (the default): tiles
 * in XCD-grouped runs of 16 on the null stream, which assumes the sort has
 * most of the GPU's workgroup slots (about 112 resident workgroups), and
 * one arrival-order ticket on any other stream, which makes progress
 * whatever else runs.  GS_TILES_ARRIVAL: the arrival-order ticket always,
 * for callers that run other work beside null-stream sorts (non-blocking
 * streams, other processes).  GS_TILES_GROUPED: grouped tiles on any
 * stream, for callers that guarantee the sort has the GPU to itself.
 * No counterpart in the reference (its seam is thrust's sort, which makes
 * no such choice visible). */
enum { GS_TILES_AUTO = -1, GS_TILES_ARRIVAL = 0, GS_TILES_GROUPED = 1 };
gs_status gs_set_tile_order(gs_handle* h, int order);
gs_status gs_synchronize(gs_handle* h);
/* Bytes currently held by the handle's arena. */
gs_status gs_arena_bytes(gs_handle* h, uint64_t* bytes);
/* Frees the arena (the next call re-grows it). */
gs_status gs_release_arena(gs_handle* h);

/* Sorts n keys of type kdt in device memory in place, ascending (descending
 * if `descending`), stream-ordered on the handle's stream.
 * Replaces the device sort of SortByThrust (src/cuda_utils.cu:403-445):
 * thrust::sort(d_v, less|greater) at :424.  Float keys: finite values only;
 * -0.0 orders before +0.0. */
gs_status gs_sort_keys(gs_handle* h, gs_dtype kdt, void* d_keys, uint64_t n, int descending);

/* Stable key-value sort of device arrays in place (payload vdt of 4 or 8
 * bytes follows its key; equal keys keep their input order).
 * Replaces thrust::sort_by_key (src/cuda_utils.cu:447-506, :477). */
gs_status gs_sort_pairs(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals, uint64_t n,
                        int descending);

/* Bytes of the fixed device array that gs_sort_pairs_noalloc uses as its
 * workspace (one per GPU, held from the first gs_create on that GPU until the
 * process ends).  It holds the workspace of a sort of 64 MiB of keys and
 * payloads: e.g. 2^24 u32 keys alone, 2^23 u32 keys with u32 payloads, 2^22
 * u64 keys with u64 payloads. */
#define GS_NOALLOC_SCRATCH_BYTES (72ull << 20)

/* gs_sort_pairs (vdt GS_NONE with d_vals NULL: gs_sort_keys) that allocates
 * nothing: no hipMalloc, hipFree or allocating library call, whatever the
 * handle's arena holds.  Its workspace is the fixed array above.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Limit: a sort whose workspace exceeds
 * GS_NOALLOC_SCRATCH_BYTES fails with GS_INVALID_ARG.  Calls share the
 * GPU's array, so calls that overlap on one GPU (other streams, threads or
 * handles) race; stream-ordered calls on one stream are safe.  Results equal
 * gs_sort_pairs / gs_sort_keys on the same input. */
gs_status gs_sort_pairs_noalloc(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals, uint64_t n,
                                int descending);

/* Host-memory forms: copy in, sort, copy out, synchronised before return.
 * gs_sort_keys_host replaces CudaUtils::Sort(comp, rank, first, last)
 * (include/cuda/cuda_utils.h:16-19, src/cuda_utils.cu:97-167).
 * gs_sort_pairs_host replaces CudaUtils::SortByKey(comp, rank, kf, kl, vf,
 * _out) (include/cuda/cuda_utils.h:21-27, src/cuda_utils.cu:169-243): keys
 * and values are sorted in place and, if h_pairs_out is not NULL, the pairs
 * are also written there as records of pair_stride bytes with the key at
 * offset 0 and the value at val_offset (the std::pair<K,V> layout of
 * src/cuda_utils.cu:498-502).
 * Inputs of more than budget / (3 * (key + value bytes)) elements are sorted
 * in chunks of that many, each on the device, then merged by a binary tree
 * of stable 2-way merges streamed through the device in blocks of a chunk
 * (the reference's SplitChunks and ParallelMergeArrays,
 * src/cuda_utils.cu:508-520, :323-371); the budget is the device's free
 * memory at the call unless gs_set_host_sort_budget set one (an input
 * whose staging, 3 * n * element bytes, is at most 256 MiB skips that query
 * and is sorted in one piece).  A chunked call also allocates host memory of
 * the input's size. */
gs_status gs_sort_keys_host(gs_handle* h, gs_dtype kdt, void* h_keys, uint64_t n, int descending);
gs_status gs_sort_pairs_host(gs_handle* h, gs_dtype kdt, void* h_keys, gs_dtype vdt, void* h_vals, uint64_t n,
                             int descending, void* h_pairs_out, uint32_t pair_stride, uint32_t val_offset);

/* Device bytes one host-memory sort of this handle may stage; 0 (the
 * default) means the device's free memory at each call, as the reference's
 * SplitChunks reads it (src/cuda_utils.cu:508-520).  A smaller budget
 * leaves room for other work on the device; a budget below three elements
 * makes host-memory sorts fail with GS_OUT_OF_MEMORY. */
gs_status gs_set_host_sort_budget(gs_handle* h, uint64_t bytes);

/* Batched segmented sort: segment s is [d_offsets[s], d_offsets[s+1]) of the
 * device arrays (offsets in device memory, nseg+1 entries, non-decreasing).
 * Every segment is sorted independently (stable with payload) in one call.
 * d_vals may be NULL with vdt GS_NONE; d_keys and d_vals may be NULL when
 * every segment is empty.  No reference counterpart as one call:
 * it batches many CudaUtils::Sort / SortByKey calls (src/cuda_utils.cu:97-243). */
gs_status gs_sort_segments(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals,
                           const uint64_t* d_offsets, uint32_t nseg, int descending);

/* 2-way merge of sorted device runs a (na) and b (nb) into out (na+nb), ties
 * from a first.  Replaces OmpUtils::Merge (include/omp/omp_utils.h:54-145). */
gs_status gs_merge(gs_handle* h, gs_dtype kdt, const void* d_a_keys, const void* d_a_vals, uint64_t na,
                   const void* d_b_keys, const void* d_b_vals, uint64_t nb, gs_dtype vdt, void* d_out_keys,
                   void* d_out_vals, int descending);

/* Bucket boundaries of nruns sorted device runs (run r = [run_off[r],
 * run_off[r+1]) of d_keys; h_run_off on the host, nruns+1 entries) cut by
 * nspl host splitters shared by all runs.  h_bounds receives nruns*(nspl+2)
 * offsets: run start, the lower_bound (upper_bound if `upper`) of each
 * splitter, run end.
 * Replaces the split loops of HykSortHelper::TransferData
 * (include/helpers/hyksort_helper.h:563-568, lower_bound) and
 * SampleSortHelper::TransferData (include/helpers/samplesort_helper.h:525-537,
 * upper_bound). */
gs_status gs_bucket_bounds(gs_handle* h, gs_dtype kdt, const void* d_keys, const uint64_t* h_run_off, uint32_t nruns,
                           const void* h_splitters, uint32_t nspl, int upper, int descending, uint64_t* h_bounds);

/* Report of one virtual-rank run (filled by gs_vrank_sort). */
typedef struct gs_vrank_report {
  uint32_t rounds;             /* exchange rounds run */
  uint32_t max_splitters;      /* capacity of `splitters` (in keys) given by the caller */
  uint32_t n_splitters;        /* splitters written */
  void* splitters;             /* caller buffer, key type: every round's splitters, group by group */
  uint64_t* bucket_sizes;      /* caller buffer: per round, per source rank, per bucket */
  uint32_t max_bucket_sizes;   /* capacity of bucket_sizes */
  uint32_t n_bucket_sizes;     /* entries written */
  uint64_t* final_rank_sizes;  /* caller buffer, p entries: keys held by each virtual rank at the end */
  double max_rank_error;       /* max |rank(splitter) - target| / N over all rounds (HykSort quality bound) */
  uint32_t splitter_iterations; /* HykSort: selection iterations, summed over rounds (largest of a round's groups) */
} gs_vrank_report;

/* Per-rank pipeline of HykSort / SampleSort / HyperQuickSort over p virtual
 * ranks held as device segments of one array.  Rank r starts with keys
 * [h_rank_off[r], h_rank_off[r+1]) (host offsets, p+1 entries).  Runs the
 * local sort of every rank, then the algorithm's rounds: splitters from
 * `samples_per_rank` regular samples, bucket bounds, the virtual exchange
 * (counts -> offsets -> addressing) and the k-way merge of received runs.
 * On return d_keys/d_vals hold the global sorted order, rank by rank; a
 * stable sort when d_vals carries the original index.
 * Replaces SortAmongProcesses (include/helpers/hyksort_helper.h:453-536,
 * include/helpers/samplesort_helper.h:420-501,
 * include/helpers/hyperquicksort_helper.h:261-346) with the MPI exchange
 * replaced by device addressing. */
gs_status gs_vrank_sort(gs_handle* h, gs_algorithm algo, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals,
                        const uint64_t* h_rank_off, uint32_t p, uint32_t k_way, uint32_t samples_per_rank,
                        int descending, gs_vrank_report* report);

/* Seeded synthetic keys written to device memory: draw i of stream `seed` is
 * splitmix64(seed + (i + index_base + 1) * 0x9E3779B97F4A7C15).
 * dist: 0 uniform u32 (high half), 1 uniform u64, 2 Zipf u32 (inverse CDF
 * over the host table h_cdf of cdf_n ranks, rank mixed by a 32-bit
 * bijection), 3 sorted u32 (i*param + draw % param), 4 all-equal u32,
 * 5 normal f32 (Box-Muller, sigma = param, -0.0 -> +0.0), 6 u32 index,
 * 7 u64 index.  No reference counterpart in the library: the reference
 * generates test data on the host in its drivers (src/test/utils.h:425-444,
 * :586-595). */
gs_status gs_generate(gs_handle* h, int dist, uint64_t seed, void* d_out, uint64_t n, uint64_t index_base,
                      double param, const double* h_cdf, uint32_t cdf_n);

/* Kernel timing: between gs_profile_begin and gs_profile_end every kernel the
 * handle launches for a sort is bracketed by events on its stream;
 * gs_profile_end synchronises and returns per-kernel aggregates.  Bytes are
 * the algorithmic HBM bytes of the launches (one read of the inputs, one
 * write of the outputs).  Counterpart of the reference's phase timers
 * (src/cuda_utils.cu:29-34, src/timer.cc:24-32). */
typedef struct gs_kernel_stat {
  char name[48];
  uint64_t launches;
  double total_ms;
  double min_ms;
  double max_ms;
  uint64_t bytes;
} gs_kernel_stat;

gs_status gs_profile_begin(gs_handle* h);
gs_status gs_profile_end(gs_handle* h, gs_kernel_stat* out, uint32_t max_out, uint32_t* n_out);

/* Size classes of the batched sort for one (key, payload) width pair:
 * segments of at most small_max keys (at most the LDS tile) are sorted by the
 * LDS bitonic kernel, those of at most medium_max by the segmented onesweep
 * with small (4K / 3K-key) tiles, the rest with the large tiles.  Kept per
 * handle; gs_autotune_segments picks them by measurement.  No counterpart in
 * the reference (one array per call, src/cuda_utils.cu:97-243). */
typedef struct gs_seg_config {
  uint32_t small_max;
  uint64_t medium_max;
  /* Tile geometry of the small-tile class: 0 = 256 threads x 16 keys (x 12
   * when a key and payload take more than 8 bytes), 1 = 256 x 8 (x 6),
   * 2 = 512 x 16 (x 12). */
  uint32_t medium_geometry;
} gs_seg_config;
gs_status gs_get_seg_config(gs_handle* h, gs_dtype kdt, gs_dtype vdt, gs_seg_config* out);
gs_status gs_set_seg_config(gs_handle* h, gs_dtype kdt, gs_dtype vdt, const gs_seg_config* cfg);

/* One measured candidate of the autotuner. */
typedef struct gs_tune_row {
  gs_seg_config config;
  double ms; /* median time of the batched sort */
} gs_tune_row;

/* Autotuner of the batched sort: builds a synthetic batch of about
 * total_keys keys in arrays with lognormal lengths (median, sigma; clipped to
 * [1, 2^22]), uniform keys and index payloads of the given types, times every
 * candidate size-class configuration (LDS-tile limit x small-tile limit x
 * small-tile geometry) with events, stores the fastest in the handle and
 * returns it (and every candidate's time in rows, up to max_rows).
 * Synchronous. */
gs_status gs_autotune_segments(gs_handle* h, gs_dtype kdt, gs_dtype vdt, uint64_t total_keys, double median,
                               double sigma, gs_seg_config* best, gs_tune_row* rows, uint32_t max_rows,
                               uint32_t* n_rows);

/* A plan of the batched sort for fixed host segment offsets, types and
 * direction (with the handle's size classes at creation): the device maps and
 * workspace are built once and owned by the plan, so running it is kernel
 * launches only -- no host copies, allocation or synchronisation -- and can
 * be captured into a HIP graph.  Segments must be shorter than 2^30 keys.
 * Runs on the handle's stream (or the plan's own stream when the handle has
 * none); calls on one plan must not overlap. */
typedef struct gs_seg_plan gs_seg_plan;
gs_status gs_seg_plan_create(gs_handle* h, gs_dtype kdt, gs_dtype vdt, const uint64_t* h_offsets, uint32_t nseg,
                             int descending, gs_seg_plan** out);
/* Queues the planned sort of d_keys / d_vals (asynchronous). */
gs_status gs_seg_plan_run(gs_seg_plan* p, void* d_keys, void* d_vals);
/* Captures the planned sort of these buffers into a HIP graph on first use
 * (again when the buffers change) and launches the graph (asynchronous). */
gs_status gs_seg_plan_launch_graph(gs_seg_plan* p, void* d_keys, void* d_vals);
gs_status gs_seg_plan_destroy(gs_seg_plan* p);

/* Stable multisplit of unsorted device keys (and payloads) into nspl + 1
 * buckets, nspl <= 255, by the sorted splitters h_splitters (host, key type,
 * sorted in the direction's order): a key's bucket is the number of
 * splitters <= key (upper = 0: the lower_bound cut of HykSort's TransferData,
 * include/helpers/hyksort_helper.h:563-568) or < key (upper = 1: the
 * upper_bound cut of SampleSort's, include/helpers/samplesort_helper.h:
 * 525-537).  On sorted input the bucket offsets equal those cuts; the kernel
 * also takes unsorted input.  d_out_keys / d_out_vals (not aliasing the
 * input) receive the buckets in order, keys of a bucket in input order;
 * h_bucket_offsets receives nspl + 2 offsets (0, ..., n).  Synchronous. */
gs_status gs_multisplit(gs_handle* h, gs_dtype kdt, const void* d_keys, gs_dtype vdt, const void* d_vals, uint64_t n,
                        const void* h_splitters, uint32_t nspl, int upper, int descending, void* d_out_keys,
                        void* d_out_vals, uint64_t* h_bucket_offsets);

/* Host-memory form of gs_vrank_sort: the arrays are staged through the
 * handle's arena and copied back in place. */
gs_status gs_vrank_sort_host(gs_handle* h, gs_algorithm algo, gs_dtype kdt, void* h_keys, gs_dtype vdt, void* h_vals,
                             const uint64_t* h_rank_off, uint32_t p, uint32_t k_way, uint32_t samples_per_rank,
                             int descending, gs_vrank_report* report);

/* dst[i] = src[idx[i]] for a payload of elem_bytes (1, 2, 4, 8) bytes.
 * Device form of ParUtils::ArrangeValues (include/par/par_utils.h:22-86). */
gs_status gs_gather(gs_handle* h, int elem_bytes, const void* d_src, const uint64_t* d_idx, void* d_dst,
                    uint64_t n);

/* Host-memory form of gs_gather: src has n_src elements, dst and idx n. */
gs_status gs_gather_host(gs_handle* h, int elem_bytes, const void* h_src, uint64_t n_src, const uint64_t* h_idx,
                         void* h_dst, uint64_t n);

/* Final rank sizes after ParUtils::KeepNoProcessHasEmptyData
 * (include/par/par_utils.h:103-175): empty ranks take data from their
 * nearest non-empty neighbour; the data keeps its global order.  Host only. */
gs_status gs_rebalance_sizes(uint64_t* sizes, uint32_t p);

/* ------------------------------------------------------------------ serving
 * A sort server: continuous batching of device-resident sort requests on one
 * GPU.  No counterpart in the reference, whose callers each run their own
 * sort call (src/cuda_utils.cu:97-243) with per-call device allocation.
 * N stream slots (each its own non-blocking stream) with pooled HBM
 * workspace; a scheduler thread admits
 * requests in arrival order when a slot is free and the batch's workspace
 * fits the HBM budget, coalesces small requests of one (key type, payload
 * type, direction) into one segmented launch, runs larger ones on their own
 * slot concurrently, and completes requests by event.  A request's buffers
 * belong to the server until it is done; its result equals gs_sort_keys /
 * gs_sort_pairs on the same input.  serial = 1 gives the one-call-per-request
 * baseline (one slot, no coalescing).  Calls are thread-safe. */
typedef struct gs_server gs_server;

typedef struct gs_server_config {
  int device;
  uint64_t hbm_budget;          /* bytes of pooled workspace the server may hold */
  uint32_t streams;             /* stream slots, 1..8 */
  uint64_t coalesce_max_keys;   /* requests of at most this many keys are coalesced */
  uint64_t batch_max_keys;      /* keys per coalesced launch (0: 2^24) */
  uint32_t batch_max_requests;  /* requests per coalesced launch (0: 4096) */
  int serial;                   /* 1: one request at a time, no coalescing */
} gs_server_config;

typedef struct gs_request_info {
  uint64_t id, n;
  gs_dtype kdt, vdt;
  gs_status status;
  int done;
  double submit_us, start_us, done_us; /* server clock: submission, launch, completion seen */
  uint64_t batch;                      /* launch the request ran in */
  uint32_t batch_size;                 /* requests in that launch */
  /* device clock (microseconds from an event taken at server creation):
   * when the launch's stream reached its start and its completion; -1 until
   * done.  Launches on different slots overlap here when their streams ran
   * side by side. */
  double dev_start_us, dev_done_us;
} gs_request_info;

typedef struct gs_server_stats {
  uint64_t submitted, completed, failed, refused;
  uint64_t batches, coalesced_batches, coalesced_requests;
  uint64_t pooled_bytes, peak_pooled_bytes, budget;
} gs_server_stats;

gs_status gs_server_create(gs_server** out, const gs_server_config* cfg);
/* Queues a request; GS_ADMISSION_REFUSED when its workspace alone exceeds the budget. */
gs_status gs_server_submit(gs_server* s, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals, uint64_t n,
                           int descending, uint64_t* id);
/* Queues a request whose keys / payloads are in host memory: they are copied
 * into the server's pooled device staging, sorted and copied back in place.
 * Pinned host memory (hipHostMalloc) lets copies overlap other requests'
 * sorts; pageable memory works with synchronous copies. */
gs_status gs_server_submit_host(gs_server* s, gs_dtype kdt, void* h_keys, gs_dtype vdt, void* h_vals, uint64_t n,
                                int descending, uint64_t* id);
/* Waits for one request (timeout_us < 0: no limit); GS_NOT_READY on timeout. */
gs_status gs_server_wait(gs_server* s, uint64_t id, int64_t timeout_us);
/* Waits for every submitted request. */
gs_status gs_server_drain(gs_server* s);
gs_status gs_server_info(gs_server* s, uint64_t id, gs_request_info* out);
gs_status gs_server_get_stats(gs_server* s, gs_server_stats* out);
/* Finishes the queued work, then frees the server. */
gs_status gs_server_destroy(gs_server* s);

#ifdef __cplusplus
}
#endif

#endif /* GPUSORT_HIP_H_ */
