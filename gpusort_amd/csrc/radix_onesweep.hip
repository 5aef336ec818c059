// Onesweep LSD radix sort for gfx950 (wave64, no MFMA).
//
// Replaces the device sort of the reference's local-sort seam:
//   thrust::sort         (src/cuda_utils.cu:424)       -- keys only
//   thrust::sort_by_key  (src/cuda_utils.cu:477-478)   -- stable, with payload
//
// One sort is 3 + P launches on one stream, with no host synchronisation:
//   fill      zero the global digit histograms
//   hist      one read of all keys -> the 256-bin histogram of every digit
//             position, and zero of every pass's look-back status words
//   scan      exclusive digit offsets per pass; passes whose digit is the same
//             for every key are marked skipped and the ping-pong buffer of
//             every remaining pass is chosen on the device
//   pass x P  onesweep pass (below)
//   copy      only if the result ended in the temporary buffer
//
// Onesweep pass, per tile of THREADS*ITEMS keys (tile order from an atomic
// ticket per workgroup class so every tile's predecessors are already
// resident):
//   1. coalesced wave-striped loads (lane-consecutive addresses per item)
//   2. stable ranking inside each wave: the lanes holding the same 8-bit
//      digit are found through a per-wave LDS lane-mask table, and a
//      per-wave LDS digit counter gives each key its rank
//   3. per-digit scan across waves and digits -> local slot of every key;
//      the tile's digit counts are published at once (decoupled look-back)
//   4. keys (and payloads, side by side) are written to LDS at their local
//      slot (tile sorted by digit)
//   5. 256 look-back lanes accumulate predecessor counts until an inclusive
//      prefix is found, then publish this tile's inclusive prefix
//   6. LDS is read back in slot order and written to global memory: lanes
//      that hold the same digit write consecutive addresses.
#include <atomic>

#include "gs_device.h"
#include "gs_internal.h"

namespace gs {

constexpr int kRadix = 256;
// Predecessor status words read per look-back step.
constexpr int kLookback = 4;

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <typename U>
struct Vec16;
template <>
struct Vec16<uint32_t> {
  using T = uint4;
  static constexpr int N = 4;
  __device__ static uint32_t get(const T& v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; }
};
template <>
struct Vec16<uint64_t> {
  using T = ulonglong2;
  static constexpr int N = 2;
  __device__ static uint64_t get(const T& v, int i) { return i == 0 ? v.x : v.y; }
};

// Digit histograms of every pass in one read of the keys (64-bit keys; the
// 32-bit keys' kernel follows).  Each lane adds
// into its own copy of the counters: copy c = lane % 32 sits in LDS bank c
// (word = (pass * 128 + digit / 2) * 32 + c), so the 64 adds of a wave fall
// two per bank whatever the digits, where one shared table takes random bank
// conflicts (about 5-way) on every add.  Two 16-bit counts share a word
// (digits 2k and 2k + 1); a count holds the keys of the 2 * 16 lanes of the
// workgroup that share its copy, at most 32 * kHistKeysPerThread < 2^16.
constexpr int kHistThreads = 1024;
constexpr uint64_t kHistKeysPerThread = 1024;

template <typename U, bool IDENT>
__global__ __launch_bounds__(kHistThreads) void radix_hist_kernel(const U* __restrict__ keys, uint64_t n, Xform<U> xf,
                                                                  int passes, unsigned long long* __restrict__ g_hist,
                                                                  uint4* __restrict__ zero_ptr, uint64_t zero_vec4) {
  constexpr int MAXP = sizeof(U);  // one 8-bit digit per key byte
  constexpr uint32_t WORDS = MAXP * (kRadix / 2) * 32;
  __shared__ __attribute__((aligned(16))) uint32_t s_cnt[WORDS];
  const uint32_t tid = threadIdx.x;
  const uint32_t copy = tid & 31u;
  uint4* s4 = reinterpret_cast<uint4*>(s_cnt);
  for (uint32_t i = tid; i < (uint32_t)passes * (kRadix / 2) * 32 / 4; i += kHistThreads) s4[i] = make_uint4(0, 0, 0, 0);
  const uint64_t gtid = (uint64_t)blockIdx.x * kHistThreads + tid;
  const uint64_t gstride = (uint64_t)gridDim.x * kHistThreads;
  for (uint64_t i = gtid; i < zero_vec4; i += gstride) zero_ptr[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();

  // Every key byte is a pass here (passes == MAXP): the pass offsets are
  // constants of the unrolled loop.
  auto count = [&](U u) {
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
      const uint32_t d = digit_of(u, p * 8);
      atomicAdd(&s_cnt[p * (kRadix / 2) * 32 + (d >> 1) * 32u + copy], 1u << ((d & 1u) << 4));
    }
  };
  using VT = typename Vec16<U>::T;
  constexpr int VN = Vec16<U>::N;
  const uint64_t nvec = n / VN;
  const VT* kv = reinterpret_cast<const VT*>(keys);
  for (uint64_t vi = gtid; vi < nvec; vi += gstride) {
    const VT v = kv[vi];
#pragma unroll
    for (int e = 0; e < VN; ++e) count(IDENT ? Vec16<U>::get(v, e) : to_ordered(Vec16<U>::get(v, e), xf));
  }
  if (blockIdx.x == 0)
    for (uint64_t i = nvec * VN + tid; i < n; i += kHistThreads) count(IDENT ? keys[i] : to_ordered(keys[i], xf));
  __syncthreads();
  // One thread per digit pair: the 32 copies summed, read in an order
  // rotated by thread so that a wave's reads spread over the banks.
  for (uint32_t w = tid; w < (uint32_t)passes * (kRadix / 2); w += kHistThreads) {
    uint32_t lo = 0, hi = 0;
#pragma unroll 8
    for (uint32_t c = 0; c < 32; ++c) {
      const uint32_t x = s_cnt[w * 32u + ((c + w) & 31u)];
      lo += x & 0xFFFFu;
      hi += x >> 16;
    }
    const uint32_t p = w / (kRadix / 2), d = (w % (kRadix / 2)) * 2;
    if (lo) atomicAdd(&g_hist[p * kRadix + d], (unsigned long long)lo);
    if (hi) atomicAdd(&g_hist[p * kRadix + d + 1], (unsigned long long)hi);
  }
}

// 32-bit keys: full 32-bit counters (no packing, so an add is a constant 1
// at an address from one bit-field extract), 32 bank copies of the four
// passes' 256 counters: 128 KB of LDS, one workgroup per CU, two 16-byte
// loads in flight per thread.
template <bool IDENT>
__global__ __launch_bounds__(kHistThreads) void radix_hist32_kernel(const uint32_t* __restrict__ keys, uint64_t n,
                                                                    Xform<uint32_t> xf,
                                                                    unsigned long long* __restrict__ g_hist,
                                                                    uint4* __restrict__ zero_ptr, uint64_t zero_vec4) {
  constexpr uint32_t WORDS = 4 * kRadix * 32;
  __shared__ __attribute__((aligned(16))) uint32_t s_cnt[WORDS];
  const uint32_t tid = threadIdx.x;
  const uint32_t copy = tid & 31u;
  uint4* s4 = reinterpret_cast<uint4*>(s_cnt);
  for (uint32_t i = tid; i < WORDS / 4; i += kHistThreads) s4[i] = make_uint4(0, 0, 0, 0);
  const uint64_t gtid = (uint64_t)blockIdx.x * kHistThreads + tid;
  const uint64_t gstride = (uint64_t)gridDim.x * kHistThreads;
  for (uint64_t i = gtid; i < zero_vec4; i += gstride) zero_ptr[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  auto count = [&](uint32_t u) {
    if (!IDENT) u = to_ordered(u, xf);
#pragma unroll
    for (int p = 0; p < 4; ++p) atomicAdd(&s_cnt[(p * kRadix + ((u >> (8 * p)) & 0xFFu)) * 32u + copy], 1u);
  };
  const uint64_t nvec = n / 4;
  const uint4* kv = reinterpret_cast<const uint4*>(keys);
  uint64_t vi = gtid;
  for (; vi + gstride < nvec; vi += 2 * gstride) {
    const uint4 a = kv[vi], b = kv[vi + gstride];
    count(a.x);
    count(a.y);
    count(a.z);
    count(a.w);
    count(b.x);
    count(b.y);
    count(b.z);
    count(b.w);
  }
  if (vi < nvec) {
    const uint4 a = kv[vi];
    count(a.x);
    count(a.y);
    count(a.z);
    count(a.w);
  }
  if (blockIdx.x == 0)
    for (uint64_t i = nvec * 4 + tid; i < n; i += kHistThreads) count(keys[i]);
  __syncthreads();
  for (uint32_t w = tid; w < 4 * kRadix; w += kHistThreads) {
    uint32_t sum = 0;
#pragma unroll 8
    for (uint32_t c = 0; c < 32; ++c) sum += s_cnt[w * 32u + ((c + w) & 31u)];
    if (sum) atomicAdd(&g_hist[w], (unsigned long long)sum);
  }
}

// Workgroups of the histogram kernel: at most kHistKeysPerThread keys per
// thread (the 16-bit count bound), and up to 512 workgroups to spread the
// read over the chip.
static uint64_t hist_blocks_for(uint64_t n) {
  const uint64_t per_block = (uint64_t)kHistThreads * kHistKeysPerThread;
  uint64_t blocks = (n + per_block - 1) / per_block;
  const uint64_t spread = (n + kHistThreads - 1) / kHistThreads;
  const uint64_t floor = spread < 512 ? spread : 512;
  if (blocks < floor) blocks = floor;
  return blocks < 1 ? 1 : blocks;
}

__global__ __launch_bounds__(256) void radix_scan_kernel(const unsigned long long* __restrict__ g_hist,
                                                         uint64_t* __restrict__ digit_base, SortCtrl* ctrl,
                                                         int passes, uint64_t n) {
  __shared__ uint64_t s_wsum[4];
  __shared__ uint32_t s_trivial[kMaxPasses];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int p = 0; p < passes; ++p) {
    const uint64_t c = g_hist[p * kRadix + tid];
    const uint64_t incl = wave_incl_scan64(c);
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint64_t add = 0;
    for (uint32_t w = 0; w < wave; ++w) add += s_wsum[w];
    digit_base[p * kRadix + tid] = incl - c + add;
    const int trivial = __syncthreads_or(c == n);
    if (tid == 0) s_trivial[p] = trivial;
    __syncthreads();
  }
  if (tid == 0) {
    uint32_t sel = 0, run = 0;
    for (int p = 0; p < kMaxPasses; ++p) {
      ctrl->tile_counter[p] = 0;
      for (int c = 0; c < 8; ++c) ctrl->class_counter[p][c] = 0;
      if (p >= passes || s_trivial[p]) {
        ctrl->src[p] = kSkipPass;
      } else {
        ctrl->src[p] = sel;
        sel ^= 1u;
        ++run;
      }
    }
    ctrl->final_sel = sel;
    ctrl->passes_run = run;
  }
}

template <typename U>
struct SplitTab;

template <typename U, typename V, typename S>
struct OnesweepArgs {
  U* keys0;
  U* keys1;
  V* vals0;
  V* vals1;
  uint64_t n;
  Xform<U> xf;
  int pass;
  int shift;
  const uint64_t* digit_base;  // 256 offsets of this pass
  S* status;                   // tiles*256 status words of this pass, zeroed
  SortCtrl* ctrl;
  uint32_t tiles;
  // Tile hand-out: 0 = one ticket counter in launch order; G > 0 = tiles in
  // groups of G consecutive ones per workgroup class.
  uint32_t group = 0;
  // Segmented mode only: tiles never straddle segments and the look-back
  // stops at the segment's first tile.
  const uint32_t* tile_seg;   // tile -> segment slot
  const uint32_t* seg_first;  // slot -> its first tile
  const uint64_t* seg_range;  // slot -> [start, end) element offsets, 2 words
  const uint32_t* seg_base;   // [slot][pass][256] digit starts within the segment
  int passes;                 // stride of seg_base
  // Multisplit mode only: the "digit" is the bucket of the key among the
  // sorted splitters, read from their bucket table (copied into LDS).
  const SplitTab<U>* split_tab;
  int upper;  // 1: bucket = #splitters < key; 0: #splitters <= key
};

// Tile of the j-th workgroup of class c (blockIdx % 8) to take a ticket.
// Tiles come in super-groups of 8*G consecutive tiles, class c taking tiles
// [c*G, c*G + G) of each; the tiles after the last whole super-group go round
// robin (tile = base + 8*i + c), which matches the number of workgroups of
// each class in a grid of `tiles` workgroups.  Within a class the map is
// increasing in j.  Workgroups are dealt to the 8 XCDs round robin by
// blockIdx, so consecutive tiles run on one XCD and the cache lines shared
// by the ends of their digit runs meet in that XCD's L2.
__device__ inline uint32_t class_tile(uint32_t c, uint32_t j, uint32_t G, uint32_t tiles) {
  const uint32_t full = tiles / (8u * G);
  if (j < full * G) return (j / G) * 8u * G + c * G + (j % G);
  return full * 8u * G + (j - full * G) * 8u + c;
}

// Multisplit bucket of an ordered key image x among nspl sorted splitter
// images s[0..nspl): #{j : s[j] < x} (upper) or #{j : s[j] <= x} (lower), a
// non-decreasing function of x.  It is read from a table of kSplitTab
// entries over the images: image x >= base falls in entry
// min((x - base) >> shift, kSplitTab - 1), images below base in entry 0,
// where base is the smallest splitter image and shift the least one that
// maps the largest splitter inside the table.  An entry holds the buckets of
// its first and last image, lo | hi << 8; the bucket of an image in it lies
// in [lo, hi] and is found there by a binary search over splitters lo..hi-1,
// which the entries that hold no splitter skip.  One LDS read per key
// instead of a descent of up to eight levels.
constexpr int kSplitTabBits = 12;
constexpr uint32_t kSplitTab = 1u << kSplitTabBits;

template <typename U>
struct SplitTab {
  U base;
  uint32_t shift;
  uint32_t nspl;
  U img[kRadix];                // splitter images, ascending; all ones past nspl
  uint16_t entry[kSplitTab];    // lo | hi << 8
};

template <typename U>
__device__ inline uint32_t count_splitters(const U* img, uint32_t lo, uint32_t hi, U x, int upper) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const U t = img[mid];
    if (upper ? (t < x) : (t <= x))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

template <typename U>
__device__ inline uint32_t tab_entry_of(U x, U base, uint32_t shift) {
  if (x < base) return 0;
  const U q = (U)(x - base) >> shift;
  return q < (U)kSplitTab ? (uint32_t)q : kSplitTab - 1;
}

// Builds the table from nspl <= 255 sorted splitters (raw keys): one
// workgroup of kSplitTabThreads threads.
constexpr int kSplitTabThreads = 1024;

template <typename U>
__global__ __launch_bounds__(kSplitTabThreads) void split_tab_kernel(const U* __restrict__ spl, uint32_t nspl,
                                                                     int upper, Xform<U> xf,
                                                                     SplitTab<U>* __restrict__ T) {
  __shared__ U s_img[kRadix];
  __shared__ U s_base;
  __shared__ uint32_t s_shift;
  const uint32_t tid = threadIdx.x;
  for (uint32_t j = tid; j < kRadix; j += kSplitTabThreads) {
    const U im = j < nspl ? to_ordered(spl[j], xf) : U(~U(0));
    s_img[j] = im;
    T->img[j] = im;
  }
  __syncthreads();
  if (tid == 0) {
    const U base = nspl ? s_img[0] : U(0);
    const U span = nspl ? (U)(s_img[nspl - 1] - base) : U(0);
    uint32_t shift = 0;
    while ((U)(span >> shift) >= (U)kSplitTab) ++shift;
    s_base = base;
    s_shift = shift;
    T->base = base;
    T->shift = shift;
    T->nspl = nspl;
  }
  __syncthreads();
  const U base = s_base;
  const uint32_t shift = s_shift;
  const U last = (U)(~U(0) - base) >> shift;  // entry of the largest image, before the clamp
  for (uint32_t i = tid; i < kSplitTab; i += kSplitTabThreads) {
    uint32_t lo, hi;
    if (i != 0 && (U)i > last) {
      lo = hi = nspl;  // no image maps here
    } else {
      const U first = i == 0 ? U(0) : (U)(base + ((U)i << shift));
      const bool to_top = i == kSplitTab - 1 || (U)i == last;
      const U end = to_top ? U(~U(0)) : (U)(base + ((U)(i + 1) << shift) - 1);
      lo = count_splitters<U>(s_img, 0, nspl, first, upper);
      hi = count_splitters<U>(s_img, lo, nspl, end, upper);
    }
    T->entry[i] = (uint16_t)(lo | (hi << 8));
  }
}

// Copies the table's entries and splitter images into LDS (`ent`: kSplitTab
// u16, `img`: kRadix images).  Every thread of the block takes part; a
// barrier must follow.
template <typename U>
__device__ inline void load_split_tab(uint16_t* ent, U* img, const SplitTab<U>* T, uint32_t tid, uint32_t nthreads) {
  const uint4* src = reinterpret_cast<const uint4*>(T->entry);
  uint4* dst = reinterpret_cast<uint4*>(ent);
  for (uint32_t i = tid; i < kSplitTab * 2 / 16; i += nthreads) dst[i] = src[i];
  for (uint32_t j = tid; j < kRadix; j += nthreads) img[j] = T->img[j];
}

// Buckets of N images at once: every entry is read first; then every key
// takes one branch-free step against the splitter at the low end of its
// range (an entry holding one splitter, the common case, is decided there,
// with no loop); entries holding several splitters are then narrowed
// together, one branch-free step per key per round (an idle key re-reads a
// valid image and keeps its range), so the LDS reads of all N keys are in
// flight together.
template <typename U, int N>
__device__ inline void tab_buckets(const uint16_t* ent, const U* img, const U (&x)[N], U base, uint32_t shift,
                                   int upper, uint32_t (&b)[N]) {
  uint32_t lo[N], hi[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t e = ent[tab_entry_of<U>(x[i], base, shift)];
    lo[i] = e & 0xFFu;
    hi[i] = e >> 8;
  }
  bool more = false;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const U t = img[lo[i]];  // lo <= nspl <= 255: a valid image even when idle
    const bool act = lo[i] < hi[i];
    const bool right = upper ? (t < x[i]) : (t <= x[i]);
    hi[i] = (act && !right) ? lo[i] : hi[i];  // every later splitter is above x too
    lo[i] = (act && right) ? lo[i] + 1 : lo[i];
    more |= lo[i] < hi[i];
  }
  while (more) {
    more = false;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const bool act = lo[i] < hi[i];
      const uint32_t mid = (lo[i] + hi[i]) >> 1;  // < 256: a valid image even when idle
      const U t = img[mid];
      const bool right = upper ? (t < x[i]) : (t <= x[i]);
      lo[i] = (act && right) ? mid + 1 : lo[i];
      hi[i] = (act && !right) ? mid : hi[i];
      more |= lo[i] < hi[i];
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) b[i] = lo[i];
}

// Digit of a key: the ordered image's byte at `shift`.  IDENT is the
// unsigned-ascending case where the image is the key itself.
template <typename U, bool IDENT>
__device__ inline uint32_t key_digit(U key, const Xform<U>& xf, int shift) {
  if constexpr (IDENT) return digit_of(key, shift);
  return digit_of(to_ordered(key, xf), shift);
}

// SEG: segmented mode (every pass runs; pass p reads buffer p & 1).  SPLIT:
// multisplit mode (one pass from buffer 0 to buffer 1, digit = bucket).  Idx
// is the type of element positions: 32 bits whenever they fit.
template <typename U, typename V, bool KV, int THREADS, int ITEMS, typename S, bool IDENT, bool SEG, typename Idx,
          bool SPLIT = false>
__global__ __launch_bounds__(THREADS) void onesweep_kernel(OnesweepArgs<U, V, S> a) {
  constexpr int WAVES = THREADS / kWave;
  constexpr uint32_t TILE = THREADS * ITEMS;
  constexpr int FLAG_SHIFT = sizeof(S) * 8 - 2;
  constexpr S FLAG_AGG = S(1) << FLAG_SHIFT;
  constexpr S FLAG_PRE = S(2) << FLAG_SHIFT;
  constexpr S VAL_MASK = FLAG_AGG - 1;
  constexpr size_t BUF_ELEM = (sizeof(U) > sizeof(V) || !KV) ? sizeof(U) : sizeof(V);
  // DUAL: keys and payloads are staged side by side (keys, then payloads),
  // and written out in one loop; the ranking tables then share the tile
  // buffer.  Geometries where the larger buffer would leave fewer workgroups
  // per CU than the 16 waves the registers allow (the 160 KB of LDS shared)
  // keep one buffer reused for the payloads.  Multisplit also keeps the
  // bucket of every slot (a byte) for the write-out.
  constexpr bool DUAL =
      KV && (1024 / THREADS) * ((size_t)TILE * (sizeof(U) + sizeof(V) + (SPLIT ? 1 : 0)) + 2048) <= 160 * 1024;
  constexpr size_t BUF_BYTES = DUAL ? (size_t)TILE * (sizeof(U) + sizeof(V)) : (size_t)TILE * BUF_ELEM;

  // s_wave_hist[w][d]: wave w's count of digit d during ranking, then the
  // tile-local slot of wave w's first key with digit d.
  __shared__ uint32_t s_wave_hist_own[DUAL ? 1 : WAVES * kRadix];
  // s_match[w][d]: lane mask of wave w's lanes holding digit d for the item
  // being ranked; every entry is back to 0 after each item.  It lives in the
  // tile buffer, free until ranking ends, which leaves LDS for larger tiles.
  constexpr bool MATCH_IN_BUF = true;
  __shared__ unsigned long long s_match_own[MATCH_IN_BUF ? 1 : WAVES * kRadix];
  __shared__ Idx s_gofs[kRadix];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_tile;
  __shared__ __attribute__((aligned(16))) unsigned char s_buf[BUF_BYTES];
  __shared__ unsigned char s_bkt[SPLIT ? TILE : 1];  // bucket of the key in each slot
  static_assert(!MATCH_IN_BUF || BUF_BYTES >= WAVES * kRadix * (sizeof(unsigned long long) + (DUAL ? 4 : 0)),
                "ranking tables fit the tile buffer");
  unsigned long long* const s_match =
      MATCH_IN_BUF ? reinterpret_cast<unsigned long long*>(s_buf) : s_match_own;
  uint32_t* const s_wave_hist =
      DUAL ? reinterpret_cast<uint32_t*>(s_buf + WAVES * kRadix * sizeof(unsigned long long)) : s_wave_hist_own;
  // Multisplit: the splitters' bucket table (entries, then splitter images)
  // sits in the tile buffer after the ranking tables until the buckets are
  // found (before ranking; the buffer is written after it).
  constexpr size_t TAB_OFF = WAVES * kRadix * (sizeof(unsigned long long) + (DUAL ? 4 : 0));
  static_assert(!SPLIT || BUF_BYTES >= TAB_OFF + kSplitTab * 2 + kRadix * sizeof(U),
                "bucket table fits the tile buffer beside the ranking tables");
  uint16_t* const s_ent = reinterpret_cast<uint16_t*>(s_buf + TAB_OFF);
  U* const s_img = reinterpret_cast<U*>(s_buf + TAB_OFF + kSplitTab * 2);
  U tab_base = U(0);
  uint32_t tab_shift = 0;
  if constexpr (SPLIT) {
    tab_base = a.split_tab->base;
    tab_shift = a.split_tab->shift;
  }

  uint32_t src;
  if constexpr (SPLIT) {
    src = 0;
  } else if constexpr (SEG) {
    src = (uint32_t)a.pass & 1u;
  } else {
    src = a.ctrl->src[a.pass];
    if (src > 1u) return;  // every key has the same digit here: the pass is the identity
  }
  const U* __restrict__ kin = src ? a.keys1 : a.keys0;
  U* __restrict__ kout = src ? a.keys0 : a.keys1;
  const V* __restrict__ vin = src ? a.vals1 : a.vals0;
  V* __restrict__ vout = src ? a.vals0 : a.vals1;
  const int shift = a.shift;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t wave = tid >> 6;

  if (tid == 0) {
    if (a.group) {
      const uint32_t c = blockIdx.x & 7u;
      s_tile = class_tile(c, atomicAdd(&a.ctrl->class_counter[a.pass][c], 1u), a.group, a.tiles);
    } else {
      s_tile = atomicAdd(&a.ctrl->tile_counter[a.pass], 1u);
    }
  }
  for (uint32_t i = tid; i < WAVES * kRadix; i += THREADS) {
    s_wave_hist[i] = 0;
    s_match[i] = 0;
  }
  if constexpr (SPLIT) load_split_tab<U>(s_ent, s_img, a.split_tab, tid, THREADS);
  lds_barrier();
  // Digit of a key: its byte at `shift`, or its bucket in multisplit mode.
  auto digit_of_key = [&](U key) -> uint32_t { return key_digit<U, IDENT>(key, a.xf, shift); };
  const uint32_t tile = s_tile;
  uint64_t tile_base, remain;
  uint32_t first = 0;       // first tile of this tile's segment
  uint64_t seg_start = 0;   // element offset of the segment
  const uint64_t* gbase64 = a.digit_base;
  const uint32_t* gbase32 = nullptr;
  if constexpr (SEG) {
    const uint32_t slot = a.tile_seg[tile];
    first = a.seg_first[slot];
    seg_start = a.seg_range[2 * slot];
    tile_base = seg_start + (uint64_t)(tile - first) * TILE;
    remain = a.seg_range[2 * slot + 1] - tile_base;
    gbase32 = a.seg_base + ((uint64_t)slot * a.passes + a.pass) * kRadix;
  } else {
    tile_base = (uint64_t)tile * TILE;
    remain = a.n - tile_base;
  }
  const uint32_t valid = remain < TILE ? (uint32_t)remain : TILE;
  // The rest is compiled twice, for full tiles and for a partial last tile
  // (of the array or of a segment): with no join between the loads and the
  // ranking, each key's wait covers that key's load only, and the payload
  // loads stay in flight during the ranking.
  auto tile_body = [&](auto full_tag) {
  constexpr bool full = decltype(full_tag)::value;

  // 1. wave-striped loads: item i of lane l of wave w is key
  //    tile_base + w*64*ITEMS + i*64 + l, so item order within a wave is key order.
  const U* __restrict__ tk = kin + tile_base;
  const V* __restrict__ tv = vin + tile_base;
  const uint32_t lofs = wave * (kWave * ITEMS) + lane;
  U k[ITEMS];
  V v[KV ? ITEMS : 1];
  if (full) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) k[i] = tk[lofs + i * kWave];
    if constexpr (KV) {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) v[i] = tv[lofs + i * kWave];
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) k[i] = (lofs + i * kWave) < valid ? tk[lofs + i * kWave] : U(0);
    if constexpr (KV) {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) v[i] = (lofs + i * kWave) < valid ? tv[lofs + i * kWave] : V(0);
    }
  }

  // Keys past the end of the array take digit 255: they are the last keys of
  // the tile in key order, so they land on the last (TILE - valid) slots.
  // Buckets (multisplit) are searched once and kept packed, 4 per register.
  uint32_t bpk[SPLIT ? (ITEMS + 3) / 4 : 1];
  if constexpr (SPLIT) {
    U img[ITEMS];
    uint32_t bk[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) img[i] = to_ordered(k[i], a.xf);
    tab_buckets<U, ITEMS>(s_ent, s_img, img, tab_base, tab_shift, a.upper, bk);
#pragma unroll
    for (int q = 0; q < (ITEMS + 3) / 4; ++q) bpk[q] = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t d = (full || (lofs + i * kWave) < valid) ? bk[i] : 255u;
      bpk[i / 4] |= d << (8 * (i % 4));
    }
  }
  auto digit_at = [&](int i) -> uint32_t {
    if constexpr (SPLIT) return (bpk[i / 4] >> (8 * (i % 4))) & 0xFFu;
    const uint32_t d = digit_of_key(k[i]);
    if (full) return d;
    return (lofs + i * kWave) < valid ? d : 255u;
  };

  // 2. stable rank inside the wave.  Lanes holding the same digit are found
  //    through LDS: each lane ORs its lane bit into its digit's mask and reads
  //    the mask back, then the mask is cleared; every matching lane then
  //    writes the same new digit count, so no lane election is needed.  The
  //    OR and the clear touch 32 bits (one bank) per lane, the read 64.
  uint32_t* wh = s_wave_hist + wave * kRadix;
  unsigned long long* wm = s_match + wave * kRadix;
  uint32_t r[ITEMS];
  // One item: peers of every lane (its digit's lanes), then the wave's
  // running count of the digit.  A lane sets and clears only its half-wave's
  // 32-bit word of the mask (lanes 0-31: low word, 32-63: high word); the
  // read takes both.  With UNIFORM, an item whose 64 keys share one digit
  // takes all lanes as peers without the table.
  auto rank_items = [&](auto uniform_tag) {
    constexpr bool UNIFORM = decltype(uniform_tag)::value;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      uint32_t d;
      if constexpr (SPLIT) {
        d = digit_at(i);
      } else {
        // The empty asm keeps the digit inside this item's iteration: hoisted
        // ahead of the uniform-wave branch (as the compiler otherwise does),
        // every digit would wait for every key and hold a register.
        U kk = k[i];
        asm volatile("" : "+v"(kk));
        d = digit_of_key(kk);
        if (!full && (lofs + i * kWave) >= valid) d = 255u;
      }
      uint64_t peers;
      if (UNIFORM && ballot(d != (uint32_t)__builtin_amdgcn_readfirstlane(d)) == 0) {
        peers = __builtin_amdgcn_read_exec();
      } else {
        uint32_t* const half = reinterpret_cast<uint32_t*>(&wm[d]) + (lane >> 5);
        __hip_atomic_fetch_or(half, 1u << (lane & 31u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __builtin_amdgcn_wave_barrier();
        peers = __hip_atomic_load(&wm[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __builtin_amdgcn_wave_barrier();
        __hip_atomic_store(half, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      const uint32_t below = count_below(peers);
      const uint32_t old = wh[d];
      __builtin_amdgcn_wave_barrier();
      wh[d] = old + (uint32_t)__popcll(peers);
      __builtin_amdgcn_wave_barrier();
      r[i] = old + below;
    }
  };
  // Runs of equal keys left by the previous pass give whole waves one digit,
  // and 64 lanes on one mask word serialise.  A wave whose first item is
  // such takes the checking loop; others keep the branch-free loop, which
  // starts as soon as their first keys arrive.
  const uint32_t d_first = digit_at(0);
  if (ballot(d_first != (uint32_t)__builtin_amdgcn_readfirstlane(d_first)) == 0)
    rank_items(std::true_type{});
  else
    rank_items(std::false_type{});
  lds_barrier();

  // 3. per digit: cross-wave totals (published at once), block scan of the
  //    totals, then the tile-local start of every digit folded into the
  //    per-wave offsets.
  uint32_t total = 0, pub = 0, incl = 0;
  if (tid < kRadix) {
#pragma unroll
    for (int w = 0; w < WAVES; ++w) total += s_wave_hist[w * kRadix + tid];
    pub = total - ((tid == kRadix - 1) ? (TILE - valid) : 0u);
    uint64_t dbase = 0;
    if (tile == first) dbase = SEG ? (uint64_t)gbase32[tid] : gbase64[tid];
    const S word = (tile == first) ? (FLAG_PRE | (S)(dbase + pub)) : (FLAG_AGG | (S)pub);
    status_store(&a.status[(uint64_t)tile * kRadix + tid], word);
    incl = wave_incl_scan(total);
    if (lane == 63) s_wsum[wave] = incl;
  }
  lds_barrier();
  uint32_t local_start = 0;
  if (tid < kRadix) {
    uint32_t add = 0;
    for (uint32_t w = 0; w < wave; ++w) add += s_wsum[w];
    local_start = incl - total + add;
    uint32_t run = local_start;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
      const uint32_t c = s_wave_hist[w * kRadix + tid];
      s_wave_hist[w * kRadix + tid] = run;
      run += c;
    }
  }
  lds_barrier();

  // 4. local slots; tile sorted by digit in LDS.
  U* s_k = reinterpret_cast<U*>(s_buf);
  V* s_vd = reinterpret_cast<V*>(s_buf + (size_t)TILE * sizeof(U));
  if constexpr (DUAL) {
    // The per-wave offsets sit in the tile buffer: all are read before any
    // key or payload is staged.
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) r[i] += wh[digit_at(i)];
    lds_barrier();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      s_k[r[i]] = k[i];
      s_vd[r[i]] = v[i];
      if constexpr (SPLIT) s_bkt[r[i]] = (unsigned char)digit_at(i);
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      r[i] += wh[digit_at(i)];
      s_k[r[i]] = k[i];
      if constexpr (SPLIT) s_bkt[r[i]] = (unsigned char)digit_at(i);
    }
  }

  // 5. decoupled look-back, one lane per digit.  Each step reads a window of
  //    kLookback predecessor words at once, then consumes them newest first:
  //    aggregates are added until an inclusive prefix closes the sum; a word
  //    not yet published ends the step and is polled again.
  if (tid < kRadix) {
    constexpr int W = kLookback;
    uint64_t excl;
    if (tile == first) {
      excl = SEG ? (uint64_t)gbase32[tid] : gbase64[tid];
    } else {
      excl = 0;
      int64_t t = (int64_t)tile - 1;  // newest predecessor not yet consumed
      bool done = false;
      while (!done) {
        S w[W];
#pragma unroll
        for (int j = 0; j < W; ++j)
          w[j] = (t - j >= (int64_t)first) ? status_load_cached(&a.status[(uint64_t)(t - j) * kRadix + tid]) : S(0);
        // A 0 may be a stale cached copy: the newest unconsumed word is
        // re-read device-coherently, so an unpublished word is always polled
        // from memory.
        if ((w[0] & ~VAL_MASK) == 0) w[0] = status_load(&a.status[(uint64_t)t * kRadix + tid]);
        int used = 0;
#pragma unroll
        for (int j = 0; j < W; ++j) {
          if (done || used < j) continue;
          const S flag = w[j] & ~VAL_MASK;
          if (flag == 0) continue;
          excl += (uint64_t)(w[j] & VAL_MASK);
          used = j + 1;
          if (flag == FLAG_PRE) done = true;
        }
        t -= used;
        if (!done && used < W) __builtin_amdgcn_s_sleep(1);
      }
      status_store(&a.status[(uint64_t)tile * kRadix + tid], (S)(FLAG_PRE | (S)(excl + pub)));
    }
    s_gofs[tid] = (Idx)(seg_start + excl - local_start);
  }
  lds_barrier();

  // 6. coalesced scatter of keys (and, staged beside them, payloads) in slot
  //    order.
  if constexpr (DUAL) {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t p = (uint32_t)j * THREADS + tid;
      if (p < valid) {
        const U key = s_k[p];
        const Idx dst = s_gofs[SPLIT ? (uint32_t)s_bkt[p] : digit_of_key(key)] + (Idx)p;
        kout[dst] = key;
        vout[dst] = s_vd[p];
      }
    }
    return;
  }
  Idx g[KV ? ITEMS : 1];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint32_t p = (uint32_t)j * THREADS + tid;
    if (p < valid) {
      const U key = s_k[p];
      const Idx dst = s_gofs[SPLIT ? (uint32_t)s_bkt[p] : digit_of_key(key)] + (Idx)p;
      kout[dst] = key;
      if constexpr (KV) g[j] = dst;
    }
  }

  // 7. payloads (geometries with one buffer) through the same slots.
  if constexpr (KV) {
    lds_barrier();
    V* s_v = reinterpret_cast<V*>(s_buf);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) s_v[r[i]] = v[i];
    lds_barrier();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t p = (uint32_t)j * THREADS + tid;
      if (p < valid) vout[g[j]] = s_v[p];
    }
  }
  };
  if (valid == TILE)
    tile_body(std::true_type{});
  else
    tile_body(std::false_type{});
}

template <typename U, typename V, bool KV>
__global__ __launch_bounds__(256) void radix_final_copy_kernel(U* __restrict__ k0, const U* __restrict__ k1,
                                                               V* __restrict__ v0, const V* __restrict__ v1,
                                                               uint64_t n, const SortCtrl* ctrl) {
  if (ctrl->final_sel == 0) return;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    k0[i] = k1[i];
    if constexpr (KV) v0[i] = v1[i];
  }
}

// ------------------------------------------------------------------ host side

template <typename U, int VB>
struct Geometry;
// Tile sizes, all 1024 threads: 32-bit keys with 32-bit payloads x 19 keys
// (19K-key tiles: fewer look-back steps per key; a key and a payload per
// slot then fill the 160 KB of LDS), keys-only 32-bit keys x 30 (the most
// that fits 128 VGPRs without spills), 64-bit keys alone x 16, 13 when key
// and payload take 12 bytes, 9 for 64-bit keys with 64-bit payloads (the
// tile of keys and payloads fits the LDS).  The ranking tables share the
// tile buffer.  Multisplit passes use SplitGeometry.
template <>
struct Geometry<uint32_t, 0> {
  static constexpr int THREADS = 1024, ITEMS = 30;
};
template <>
struct Geometry<uint32_t, 4> {
  static constexpr int THREADS = 1024, ITEMS = 19;
};
template <>
struct Geometry<uint32_t, 8> {
  static constexpr int THREADS = 1024, ITEMS = 13;
};
template <>
struct Geometry<uint64_t, 0> {
  static constexpr int THREADS = 1024, ITEMS = 16;
};
template <>
struct Geometry<uint64_t, 4> {
  static constexpr int THREADS = 1024, ITEMS = 13;
};
template <>
struct Geometry<uint64_t, 8> {
  static constexpr int THREADS = 1024, ITEMS = 9;
};

// Smaller tiles for mid-size single arrays, where the large tile leaves CUs
// idle (2^22 keys-only 32-bit keys are 137 tiles of 30 K for 256 CUs):
// keys-only 32-bit keys up to 2^23 keys at 1024 x 16, 32-bit keys with
// 32-bit payloads up to 2^21 keys at 1024 x 10, keys-only 64-bit keys up to
// 2^21 keys at 1024 x 8.  MAX_N 0: no smaller tile.
template <typename U, int VB>
struct SmallGeometry {
  static constexpr int THREADS = Geometry<U, VB>::THREADS, ITEMS = Geometry<U, VB>::ITEMS;
  static constexpr uint64_t MAX_N = 0;
};
template <>
struct SmallGeometry<uint32_t, 0> {
  static constexpr int THREADS = 1024, ITEMS = 16;
  static constexpr uint64_t MAX_N = 1ull << 23;
};
template <>
struct SmallGeometry<uint32_t, 4> {
  static constexpr int THREADS = 1024, ITEMS = 10;
  static constexpr uint64_t MAX_N = 1ull << 21;
};
template <>
struct SmallGeometry<uint64_t, 0> {
  static constexpr int THREADS = 1024, ITEMS = 8;
  static constexpr uint64_t MAX_N = 1ull << 21;
};

template <int VB>
struct ValT {
  using T = uint32_t;
};
template <>
struct ValT<8> {
  using T = uint64_t;
};

static inline uint32_t tile_keys_for(int kb, int vb) {
  if (kb == 4) {
    if (vb == 0) return Geometry<uint32_t, 0>::THREADS * Geometry<uint32_t, 0>::ITEMS;
    if (vb == 4) return Geometry<uint32_t, 4>::THREADS * Geometry<uint32_t, 4>::ITEMS;
    return Geometry<uint32_t, 8>::THREADS * Geometry<uint32_t, 8>::ITEMS;
  }
  if (vb == 0) return Geometry<uint64_t, 0>::THREADS * Geometry<uint64_t, 0>::ITEMS;
  if (vb == 4) return Geometry<uint64_t, 4>::THREADS * Geometry<uint64_t, 4>::ITEMS;
  return Geometry<uint64_t, 8>::THREADS * Geometry<uint64_t, 8>::ITEMS;
}


// Multisplit passes keep a bucket byte per slot beside the side-by-side
// tile: 32-bit keys alone at 1024 x 22 (registers), with 32-bit payloads at
// 1024 x 17, 64-bit keys alone at 1024 x 14, and 12-byte key + payload at
// 1024 x 11 (64-bit keys with 64-bit payloads as Geometry, 1024 x 9).
template <typename U, int VB>
struct SplitGeometry {
  static constexpr int THREADS = Geometry<U, VB>::THREADS, ITEMS = Geometry<U, VB>::ITEMS;
};
template <>
struct SplitGeometry<uint32_t, 0> {
  static constexpr int THREADS = 1024, ITEMS = 22;
};
template <>
struct SplitGeometry<uint32_t, 4> {
  static constexpr int THREADS = 1024, ITEMS = 17;
};
template <>
struct SplitGeometry<uint32_t, 8> {
  static constexpr int THREADS = 1024, ITEMS = 11;
};
template <>
struct SplitGeometry<uint64_t, 0> {
  static constexpr int THREADS = 1024, ITEMS = 14;
};
template <>
struct SplitGeometry<uint64_t, 4> {
  static constexpr int THREADS = 1024, ITEMS = 11;
};

static inline uint32_t split_tile_keys(int kb, int vb) {
  if (kb == 4 && vb == 0) return SplitGeometry<uint32_t, 0>::THREADS * SplitGeometry<uint32_t, 0>::ITEMS;
  if (kb == 4 && vb == 4) return SplitGeometry<uint32_t, 4>::THREADS * SplitGeometry<uint32_t, 4>::ITEMS;
  if (kb == 4 && vb == 8) return SplitGeometry<uint32_t, 8>::THREADS * SplitGeometry<uint32_t, 8>::ITEMS;
  if (kb == 8 && vb == 0) return SplitGeometry<uint64_t, 0>::THREADS * SplitGeometry<uint64_t, 0>::ITEMS;
  if (kb == 8 && vb == 4) return SplitGeometry<uint64_t, 4>::THREADS * SplitGeometry<uint64_t, 4>::ITEMS;
  return SplitGeometry<uint64_t, 8>::THREADS * SplitGeometry<uint64_t, 8>::ITEMS;
}

// Tile geometries of the segmented sort: 0 = the single-array geometry above
// (keys-only 32-bit keys at 1024 x 28: the segmented pass spills at 30),
// 1 = 256 threads (4K keys for 32-bit keys with 0/4-byte payloads, else 3K),
// for segments of a few tiles: less of the last tile is padding and up to
// four workgroups share a CU.  Geometry 1 has the smallest tiles.
template <typename U, int VB, int G>
struct SegGeometry {
  static constexpr int THREADS = Geometry<U, VB>::THREADS, ITEMS = Geometry<U, VB>::ITEMS;
};
template <>
struct SegGeometry<uint32_t, 0, 0> {
  static constexpr int THREADS = 1024, ITEMS = 28;
};
template <typename U, int VB>
struct SegGeometry<U, VB, 1> {
  static constexpr int THREADS = 256, ITEMS = (sizeof(U) == 4 && VB <= 4) ? 16 : 12;
};
// The medium class's other choices for the autotuner: half the keys per
// thread (2: 256 threads x 8 / x 6), or twice the threads (3: 512 x 16 /
// x 12).
template <typename U, int VB>
struct SegGeometry<U, VB, 2> {
  static constexpr int THREADS = 256, ITEMS = SegGeometry<U, VB, 1>::ITEMS / 2;
};
template <typename U, int VB>
struct SegGeometry<U, VB, 3> {
  static constexpr int THREADS = 512, ITEMS = SegGeometry<U, VB, 1>::ITEMS;
};

uint32_t radix_seg_tile_keys(int kb, int vb, int geo) {
  if (geo == 0) {
    if (kb == 4 && vb == 0) return SegGeometry<uint32_t, 0, 0>::THREADS * SegGeometry<uint32_t, 0, 0>::ITEMS;
    return tile_keys_for(kb, vb);
  }
  const uint32_t items = (kb == 4 && vb <= 4) ? 16 : 12;
  if (geo == 2) return 256 * (items / 2);
  if (geo == 3) return 512 * items;
  return 256 * items;
}

struct RadixLayout {
  size_t off_keys_tmp, off_vals_tmp, off_ctrl, off_hist, off_base, off_status, total;
  uint64_t tiles;
  int passes;
  size_t status_word;
};

// Tile of a single-array sort of n keys (SmallGeometry up to its MAX_N).
static inline uint32_t sort_tile_keys(int kb, int vb, uint64_t n) {
  if (kb == 4 && vb == 0 && n <= SmallGeometry<uint32_t, 0>::MAX_N)
    return SmallGeometry<uint32_t, 0>::THREADS * SmallGeometry<uint32_t, 0>::ITEMS;
  if (kb == 4 && vb == 4 && n <= SmallGeometry<uint32_t, 4>::MAX_N)
    return SmallGeometry<uint32_t, 4>::THREADS * SmallGeometry<uint32_t, 4>::ITEMS;
  if (kb == 8 && vb == 0 && n <= SmallGeometry<uint64_t, 0>::MAX_N)
    return SmallGeometry<uint64_t, 0>::THREADS * SmallGeometry<uint64_t, 0>::ITEMS;
  return tile_keys_for(kb, vb);
}

static RadixLayout radix_layout(int kb, int vb, uint64_t n) {
  RadixLayout L{};
  const uint32_t tile = sort_tile_keys(kb, vb, n);
  L.tiles = (n + tile - 1) / tile;
  L.passes = kb;  // 8-bit digits: one pass per key byte
  L.status_word = (n < (1ull << 30)) ? 4 : 8;
  size_t o = 0;
  L.off_keys_tmp = o;
  o = align_up(o + n * kb, 256);
  L.off_vals_tmp = o;
  o = align_up(o + n * vb, 256);
  L.off_ctrl = o;
  o = align_up(o + sizeof(SortCtrl), 256);
  L.off_hist = o;
  o = align_up(o + kMaxPasses * kRadix * sizeof(unsigned long long), 256);
  L.off_base = o;
  o = align_up(o + kMaxPasses * kRadix * sizeof(uint64_t), 256);
  L.off_status = o;
  o = align_up(o + (size_t)L.passes * L.tiles * kRadix * L.status_word, 256);
  L.total = o;
  return L;
}

size_t radix_workspace_bytes(int key_bytes, int val_bytes, uint64_t n) {
  return radix_layout(key_bytes, val_bytes, n).total;
}

template <typename U>
static Xform<U> xform_of(const XformParams& p) {
  return Xform<U>{(U)p.float_mask, (U)p.sign_flip, (U)p.desc_mask};
}

// Tiles per XCD group of a pass.  A tile's look-back waits on every lower
// tile.  With one ticket counter taken in arrival order (G = 0) every lower
// tile is held by a workgroup that has already started, so the pass finishes
// whatever else runs on the GPU.  With grouped hand-out the lower tiles of a
// super-group may belong to workgroups dispatched up to 8*G positions later,
// so progress needs about 8*G resident workgroups of this kernel: true when it
// has the GPU to itself (at least one per CU), not when other spinning
// kernels (another sort on another stream) can hold the CUs.  Grouping is
// therefore used only when the caller says the pass has the GPU to itself
// (kTilesGrouped), or by default on the null stream, which the library's own
// concurrent users (server slots, handles given a stream) never use: G is the
// largest power of two up to 16 with 16*G <= CUs, 0 below 16 CUs, on any
// other stream by default, or with kTilesArrival.
static uint32_t tile_group(hipStream_t stream, int policy) {
  if (policy == kTilesArrival || (policy == kTilesByStream && stream != nullptr)) return 0;
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  int known = cache[dev].load(std::memory_order_relaxed);
  if (known == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    int g = 0;
    for (int c = 1; c <= 16; c <<= 1)
      if (16 * c <= cus) g = c;
    known = g + 1;  // 0 = not yet known
    cache[dev].store(known, std::memory_order_relaxed);
  }
  return (uint32_t)(known - 1);
}

template <typename U, int VB, typename S, int THREADS, int ITEMS>
static hipError_t radix_sort_t(U* keys, void* vals_v, uint64_t n, const XformParams& xp, unsigned char* ws,
                               const RadixLayout& L, hipStream_t stream, LaunchTimer* timer, int policy) {
  using V = typename ValT<VB>::T;
  // Positions fit 32 bits whenever the 32-bit status word is used (n < 2^30).
  using Idx = typename std::conditional<sizeof(S) == 4, uint32_t, uint64_t>::type;
  constexpr bool KV = VB != 0;
  // The layout (status words, grid) and this launch geometry must agree on
  // the tile, or the passes would index status words past their region.
  if (L.tiles != (n + (uint64_t)THREADS * ITEMS - 1) / ((uint64_t)THREADS * ITEMS)) return hipErrorInvalidValue;
  const Xform<U> xf = xform_of<U>(xp);
  const bool ident = xf.float_mask == 0 && xf.sign_flip == 0 && xf.desc_mask == 0;
  U* keys_tmp = reinterpret_cast<U*>(ws + L.off_keys_tmp);
  V* vals = reinterpret_cast<V*>(vals_v);
  V* vals_tmp = reinterpret_cast<V*>(ws + L.off_vals_tmp);
  SortCtrl* ctrl = reinterpret_cast<SortCtrl*>(ws + L.off_ctrl);
  unsigned long long* hist = reinterpret_cast<unsigned long long*>(ws + L.off_hist);
  uint64_t* base = reinterpret_cast<uint64_t*>(ws + L.off_base);
  S* status = reinterpret_cast<S*>(ws + L.off_status);
  const uint64_t status_per_pass = L.tiles * kRadix;

  hipLaunchKernelGGL(fill_u32_kernel, dim3(8), dim3(256), 0, stream, reinterpret_cast<uint32_t*>(hist),
                     (uint64_t)(kMaxPasses * kRadix * 2), 0u);
  const uint64_t status_bytes = (uint64_t)L.passes * status_per_pass * sizeof(S);
  const uint64_t hist_blocks = hist_blocks_for(n);
  // The histogram kernel counts every key byte (L.passes == sizeof(U)).
  if (timer) timer->begin("radix_hist", n * sizeof(U) + status_bytes, stream);
  if constexpr (sizeof(U) == 4) {
    // One workgroup per CU at most (128 KB of LDS each); fewer for small n.
    const uint32_t blocks32 = (uint32_t)std::min<uint64_t>(256, hist_blocks);
    if (ident)
      hipLaunchKernelGGL((radix_hist32_kernel<true>), dim3(blocks32), dim3(kHistThreads), 0, stream,
                         reinterpret_cast<const uint32_t*>(keys), n, *reinterpret_cast<const Xform<uint32_t>*>(&xf),
                         hist, reinterpret_cast<uint4*>(status), status_bytes / 16);
    else
      hipLaunchKernelGGL((radix_hist32_kernel<false>), dim3(blocks32), dim3(kHistThreads), 0, stream,
                         reinterpret_cast<const uint32_t*>(keys), n, *reinterpret_cast<const Xform<uint32_t>*>(&xf),
                         hist, reinterpret_cast<uint4*>(status), status_bytes / 16);
  } else if (ident)
    hipLaunchKernelGGL((radix_hist_kernel<U, true>), dim3((uint32_t)hist_blocks), dim3(kHistThreads), 0, stream, keys,
                       n, xf, L.passes, hist, reinterpret_cast<uint4*>(status), status_bytes / 16);
  else
    hipLaunchKernelGGL((radix_hist_kernel<U, false>), dim3((uint32_t)hist_blocks), dim3(kHistThreads), 0, stream,
                       keys, n, xf, L.passes, hist, reinterpret_cast<uint4*>(status), status_bytes / 16);
  if (timer) timer->end(stream);
  hipLaunchKernelGGL(radix_scan_kernel, dim3(1), dim3(256), 0, stream, hist, base, ctrl, L.passes, n);
  for (int p = 0; p < L.passes; ++p) {
    OnesweepArgs<U, V, S> a;
    a.keys0 = keys;
    a.keys1 = keys_tmp;
    a.vals0 = vals;
    a.vals1 = vals_tmp;
    a.n = n;
    a.xf = xf;
    a.pass = p;
    a.shift = 8 * p;
    a.digit_base = base + p * kRadix;
    a.status = status + p * status_per_pass;
    a.ctrl = ctrl;
    a.tiles = (uint32_t)L.tiles;
    a.group = tile_group(stream, policy);
    if (timer) timer->begin("onesweep_pass", 2 * n * (sizeof(U) + (KV ? sizeof(V) : 0)), stream);
    if (ident)
      hipLaunchKernelGGL((onesweep_kernel<U, V, KV, THREADS, ITEMS, S, true, false, Idx>), dim3((uint32_t)L.tiles),
                         dim3(THREADS), 0, stream, a);
    else
      hipLaunchKernelGGL((onesweep_kernel<U, V, KV, THREADS, ITEMS, S, false, false, Idx>), dim3((uint32_t)L.tiles),
                         dim3(THREADS), 0, stream, a);
    if (timer) timer->end(stream);
  }
  uint64_t copy_blocks = (n + 255) / 256;
  if (copy_blocks > 4096) copy_blocks = 4096;
  hipLaunchKernelGGL((radix_final_copy_kernel<U, V, KV>), dim3((uint32_t)copy_blocks), dim3(256), 0, stream, keys,
                     keys_tmp, vals, vals_tmp, n, ctrl);
  return hipGetLastError();
}

template <typename U, int VB>
static hipError_t radix_sort_s(U* keys, void* vals, uint64_t n, const XformParams& xp, unsigned char* ws,
                               const RadixLayout& L, hipStream_t stream, LaunchTimer* timer, int policy) {
  using G = Geometry<U, VB>;
  using SG = SmallGeometry<U, VB>;
  if (n <= SG::MAX_N)  // n < 2^30: 32-bit status words
    return radix_sort_t<U, VB, uint32_t, SG::THREADS, SG::ITEMS>(keys, vals, n, xp, ws, L, stream, timer, policy);
  if (L.status_word == 4)
    return radix_sort_t<U, VB, uint32_t, G::THREADS, G::ITEMS>(keys, vals, n, xp, ws, L, stream, timer, policy);
  return radix_sort_t<U, VB, uint64_t, G::THREADS, G::ITEMS>(keys, vals, n, xp, ws, L, stream, timer, policy);
}

hipError_t radix_sort(int kb, int vb, void* keys, void* vals, uint64_t n, const XformParams& xf, void* workspace,
                      size_t workspace_bytes, hipStream_t stream, LaunchTimer* timer, int policy) {
  if (n == 0) return hipSuccess;
  const RadixLayout L = radix_layout(kb, vb, n);
  if (workspace_bytes < L.total) return hipErrorInvalidValue;
  unsigned char* ws = reinterpret_cast<unsigned char*>(workspace);
  if (kb == 4) {
    uint32_t* k = reinterpret_cast<uint32_t*>(keys);
    if (vb == 0) return radix_sort_s<uint32_t, 0>(k, vals, n, xf, ws, L, stream, timer, policy);
    if (vb == 4) return radix_sort_s<uint32_t, 4>(k, vals, n, xf, ws, L, stream, timer, policy);
    if (vb == 8) return radix_sort_s<uint32_t, 8>(k, vals, n, xf, ws, L, stream, timer, policy);
  } else if (kb == 8) {
    uint64_t* k = reinterpret_cast<uint64_t*>(keys);
    if (vb == 0) return radix_sort_s<uint64_t, 0>(k, vals, n, xf, ws, L, stream, timer, policy);
    if (vb == 4) return radix_sort_s<uint64_t, 4>(k, vals, n, xf, ws, L, stream, timer, policy);
    if (vb == 8) return radix_sort_s<uint64_t, 8>(k, vals, n, xf, ws, L, stream, timer, policy);
  }
  return hipErrorInvalidValue;
}

// ------------------------------------------------------------ segmented sort
//
// Many segments sorted by one launch sequence: fill, seg_hist, seg_scan, then
// one onesweep launch per digit for all of them.  Each segment is cut into
// its own tiles (the last one partial), tiles are numbered segment after
// segment, and a tile's look-back ends at its segment's first tile, whose
// prefix is the segment's digit start.  Segments must be shorter than 2^30
// keys (32-bit look-back words).

template <typename U>
struct SegHistArgs {
  const U* keys;
  Xform<U> xf;
  int passes;
  uint32_t tile_keys;
  const uint32_t* tile_seg;
  const uint32_t* seg_first;
  const uint64_t* seg_range;
  uint32_t n_multi;  // slots [0, n_multi) have several tiles
  uint32_t* seg_cnt;  // [slot][pass][256]
  uint32_t* status;   // [pass][tile][256]
  uint64_t status_per_pass;
};

// Digit counts of one tile for every pass, added into its segment's counts
// (atomically for multi-tile segments, whose counts start at zero; stored
// for single-tile ones); also zeroes the tile's look-back words.  Counters
// come in kSegHistCopies copies interleaved by lane (lane % 4), so lanes
// with one digit (the zero high bytes of small 64-bit keys) add into 16-way
// rather than 64-way conflicting words; a key that the whole wave holds
// (runs of equal keys) is added once per digit by one lane.
constexpr uint32_t kSegHistCopies = 4;

template <typename U, bool IDENT>
__global__ __launch_bounds__(256) void seg_hist_kernel(SegHistArgs<U> a) {
  constexpr int MAXP = sizeof(U);  // one 8-bit digit per key byte (passes == MAXP)
  __shared__ uint32_t s_hist[MAXP * kRadix * kSegHistCopies];  // [pass][digit][copy]
  const uint32_t tid = threadIdx.x, tile = blockIdx.x;
  const uint32_t copy = tid & (kSegHistCopies - 1);
  const int passes = a.passes;
  for (uint32_t i = tid; i < (uint32_t)MAXP * kRadix * kSegHistCopies; i += 256) s_hist[i] = 0;
  for (int p = 0; p < passes; ++p) a.status[p * a.status_per_pass + (uint64_t)tile * kRadix + tid] = 0;
  const uint32_t slot = a.tile_seg[tile];
  const uint64_t start = a.seg_range[2 * slot] + (uint64_t)(tile - a.seg_first[slot]) * a.tile_keys;
  const uint64_t seg_end = a.seg_range[2 * slot + 1];
  const uint64_t end = (start + a.tile_keys < seg_end) ? start + a.tile_keys : seg_end;
  __syncthreads();
  auto count = [&](U key) {
    const U u = IDENT ? key : to_ordered(key, a.xf);
    const uint64_t active = __builtin_amdgcn_read_exec();
    U u0;
    if constexpr (sizeof(U) == 8)
      u0 = ((U)__builtin_amdgcn_readfirstlane((uint32_t)(u >> 32)) << 32) | __builtin_amdgcn_readfirstlane((uint32_t)u);
    else
      u0 = __builtin_amdgcn_readfirstlane(u);
    if (ballot(u == u0) == active) {
      if (count_below(active) == 0) {
#pragma unroll
        for (int p = 0; p < MAXP; ++p)
          atomicAdd(&s_hist[(p * kRadix + digit_of(u, p * 8)) * kSegHistCopies], (uint32_t)__popcll(active));
      }
    } else {
#pragma unroll
      for (int p = 0; p < MAXP; ++p) atomicAdd(&s_hist[(p * kRadix + digit_of(u, p * 8)) * kSegHistCopies + copy], 1u);
    }
  };
  // Four keys per thread in flight: the loads are issued before any is
  // counted.
  uint64_t i = start + tid;
  for (; i + 3 * 256 < end; i += 4 * 256) {
    const U k0 = a.keys[i], k1 = a.keys[i + 256], k2 = a.keys[i + 512], k3 = a.keys[i + 768];
    count(k0);
    count(k1);
    count(k2);
    count(k3);
  }
  for (; i < end; i += 256) count(a.keys[i]);
  __syncthreads();
  uint32_t* dst = a.seg_cnt + (uint64_t)slot * passes * kRadix;
  for (uint32_t i = tid; i < (uint32_t)passes * kRadix; i += 256) {
    uint32_t c = 0;
#pragma unroll
    for (uint32_t j = 0; j < kSegHistCopies; ++j) c += s_hist[i * kSegHistCopies + j];
    if (slot < a.n_multi) {
      if (c) atomicAdd(&dst[i], c);
    } else {
      dst[i] = c;
    }
  }
}

// Counts -> exclusive digit starts within the segment, in place, one
// workgroup per segment; block 0 also resets the pass tickets.
__global__ __launch_bounds__(256) void seg_scan_kernel(uint32_t* seg_cnt, int passes, SortCtrl* ctrl) {
  __shared__ uint32_t s_wsum[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t* c = seg_cnt + (uint64_t)blockIdx.x * passes * kRadix;
  for (int p = 0; p < passes; ++p) {
    const uint32_t x = c[p * kRadix + tid];
    const uint32_t incl = wave_incl_scan(x);
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t w = 0; w < wave; ++w) add += s_wsum[w];
    c[p * kRadix + tid] = incl - x + add;
    __syncthreads();
  }
  if (blockIdx.x == 0 && tid < kMaxPasses) ctrl->tile_counter[tid] = 0;
  if (blockIdx.x == 0 && tid < kMaxPasses * 8) ctrl->class_counter[tid / 8][tid % 8] = 0;
}

struct SegLayout {
  size_t off_keys_tmp, off_vals_tmp, off_ctrl, off_cnt, off_status, total;
  int passes;
};

static SegLayout seg_layout(int kb, int vb, uint64_t n_total, uint64_t tiles, uint32_t nslots) {
  SegLayout L{};
  L.passes = kb;
  size_t o = 0;
  L.off_keys_tmp = o;
  o = align_up(o + n_total * kb, 256);
  L.off_vals_tmp = o;
  o = align_up(o + n_total * vb, 256);
  L.off_ctrl = o;
  o = align_up(o + sizeof(SortCtrl), 256);
  L.off_cnt = o;
  o = align_up(o + (size_t)nslots * L.passes * kRadix * 4, 256);
  L.off_status = o;
  o = align_up(o + (size_t)L.passes * tiles * kRadix * 4, 256);
  L.total = o;
  return L;
}

size_t radix_segments_workspace_bytes(int key_bytes, int val_bytes, uint64_t n_total, uint64_t tiles,
                                      uint32_t nslots) {
  return seg_layout(key_bytes, val_bytes, n_total, tiles, nslots).total;
}

template <typename U, int VB, typename Idx, int G>
static hipError_t radix_segments_t(U* keys, void* vals_v, const SegTiles& T, const XformParams& xp,
                                   unsigned char* ws, const SegLayout& L, hipStream_t stream, LaunchTimer* timer,
                                   int policy) {
  using V = typename ValT<VB>::T;
  constexpr bool KV = VB != 0;
  constexpr int THREADS = SegGeometry<U, VB, G>::THREADS;
  constexpr int ITEMS = SegGeometry<U, VB, G>::ITEMS;
  const Xform<U> xf = xform_of<U>(xp);
  const bool ident = xf.float_mask == 0 && xf.sign_flip == 0 && xf.desc_mask == 0;
  SortCtrl* ctrl = reinterpret_cast<SortCtrl*>(ws + L.off_ctrl);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(ws + L.off_cnt);
  uint32_t* status = reinterpret_cast<uint32_t*>(ws + L.off_status);
  const uint64_t status_per_pass = T.tiles * kRadix;
  // The plan's tiles were cut with radix_seg_tile_keys: it must be this
  // geometry's tile, or the tile maps and status words would not fit it.
  if (radix_seg_tile_keys((int)sizeof(U), VB, G) != (uint32_t)(THREADS * ITEMS)) return hipErrorInvalidValue;

  if (T.n_multi) {
    hipError_t e = fill_u32(cnt, (uint64_t)T.n_multi * L.passes * kRadix, 0u, stream);
    if (e != hipSuccess) return e;
  }
  SegHistArgs<U> h;
  h.keys = keys;
  h.xf = xf;
  h.passes = L.passes;
  h.tile_keys = THREADS * ITEMS;
  h.tile_seg = T.tile_seg;
  h.seg_first = T.seg_first;
  h.seg_range = T.seg_range;
  h.n_multi = T.n_multi;
  h.seg_cnt = cnt;
  h.status = status;
  h.status_per_pass = status_per_pass;
  if (timer) timer->begin("seg_hist", T.n_keys * sizeof(U), stream);
  if (ident)
    hipLaunchKernelGGL((seg_hist_kernel<U, true>), dim3((uint32_t)T.tiles), dim3(256), 0, stream, h);
  else
    hipLaunchKernelGGL((seg_hist_kernel<U, false>), dim3((uint32_t)T.tiles), dim3(256), 0, stream, h);
  if (timer) timer->end(stream);
  hipLaunchKernelGGL(seg_scan_kernel, dim3(T.nslots), dim3(256), 0, stream, cnt, L.passes, ctrl);
  for (int p = 0; p < L.passes; ++p) {
    OnesweepArgs<U, V, uint32_t> a{};
    a.keys0 = keys;
    a.keys1 = reinterpret_cast<U*>(ws + L.off_keys_tmp);
    a.vals0 = reinterpret_cast<V*>(vals_v);
    a.vals1 = reinterpret_cast<V*>(ws + L.off_vals_tmp);
    a.n = 0;
    a.xf = xf;
    a.pass = p;
    a.shift = 8 * p;
    a.digit_base = nullptr;
    a.status = status + p * status_per_pass;
    a.ctrl = ctrl;
    a.tiles = (uint32_t)T.tiles;
    a.tile_seg = T.tile_seg;
    a.seg_first = T.seg_first;
    a.seg_range = T.seg_range;
    a.seg_base = cnt;
    a.passes = L.passes;
    a.group = tile_group(stream, policy);
    if (timer) timer->begin("seg_onesweep_pass", 2 * T.n_keys * (sizeof(U) + (KV ? sizeof(V) : 0)), stream);
    if (ident)
      hipLaunchKernelGGL((onesweep_kernel<U, V, KV, THREADS, ITEMS, uint32_t, true, true, Idx>),
                         dim3((uint32_t)T.tiles), dim3(THREADS), 0, stream, a);
    else
      hipLaunchKernelGGL((onesweep_kernel<U, V, KV, THREADS, ITEMS, uint32_t, false, true, Idx>),
                         dim3((uint32_t)T.tiles), dim3(THREADS), 0, stream, a);
    if (timer) timer->end(stream);
  }
  return hipGetLastError();
}

template <typename U, int VB>
static hipError_t radix_segments_i(U* keys, void* vals, const SegTiles& T, const XformParams& xp, unsigned char* ws,
                                   const SegLayout& L, hipStream_t stream, LaunchTimer* timer, int policy) {
  const bool small_idx = T.n_total < (1ull << 32);
  if (T.geo == 1)
    return small_idx ? radix_segments_t<U, VB, uint32_t, 1>(keys, vals, T, xp, ws, L, stream, timer, policy)
                     : radix_segments_t<U, VB, uint64_t, 1>(keys, vals, T, xp, ws, L, stream, timer, policy);
  if (T.geo == 2)
    return small_idx ? radix_segments_t<U, VB, uint32_t, 2>(keys, vals, T, xp, ws, L, stream, timer, policy)
                     : radix_segments_t<U, VB, uint64_t, 2>(keys, vals, T, xp, ws, L, stream, timer, policy);
  if (T.geo == 3)
    return small_idx ? radix_segments_t<U, VB, uint32_t, 3>(keys, vals, T, xp, ws, L, stream, timer, policy)
                     : radix_segments_t<U, VB, uint64_t, 3>(keys, vals, T, xp, ws, L, stream, timer, policy);
  if (T.geo != 0) return hipErrorInvalidValue;
  if (T.n_total < (1ull << 32)) return radix_segments_t<U, VB, uint32_t, 0>(keys, vals, T, xp, ws, L, stream, timer, policy);
  return radix_segments_t<U, VB, uint64_t, 0>(keys, vals, T, xp, ws, L, stream, timer, policy);
}

hipError_t radix_sort_segments(int kb, int vb, void* keys, void* vals, const SegTiles& T, const XformParams& xf,
                               void* workspace, size_t workspace_bytes, hipStream_t stream, LaunchTimer* timer,
                               int policy) {
  if (T.tiles == 0) return hipSuccess;
  if (T.tiles >= (1ull << 31)) return hipErrorInvalidValue;
  const SegLayout L = seg_layout(kb, vb, T.n_total, T.tiles, T.nslots);
  if (workspace_bytes < L.total) return hipErrorInvalidValue;
  unsigned char* ws = reinterpret_cast<unsigned char*>(workspace);
  if (kb == 4) {
    uint32_t* k = reinterpret_cast<uint32_t*>(keys);
    if (vb == 0) return radix_segments_i<uint32_t, 0>(k, vals, T, xf, ws, L, stream, timer, policy);
    if (vb == 4) return radix_segments_i<uint32_t, 4>(k, vals, T, xf, ws, L, stream, timer, policy);
    if (vb == 8) return radix_segments_i<uint32_t, 8>(k, vals, T, xf, ws, L, stream, timer, policy);
  } else if (kb == 8) {
    uint64_t* k = reinterpret_cast<uint64_t*>(keys);
    if (vb == 0) return radix_segments_i<uint64_t, 0>(k, vals, T, xf, ws, L, stream, timer, policy);
    if (vb == 4) return radix_segments_i<uint64_t, 4>(k, vals, T, xf, ws, L, stream, timer, policy);
    if (vb == 8) return radix_segments_i<uint64_t, 8>(k, vals, T, xf, ws, L, stream, timer, policy);
  }
  return hipErrorInvalidValue;
}

// ------------------------------------------------------------------ multisplit
//
// Stable multisplit of unsorted keys (and payloads) into nspl + 1 <= 256
// buckets: the bucket table (split_tab_kernel), one counting read
// (split_hist_kernel), the scan, and one onesweep pass whose digit is the
// key's bucket, read from the table in LDS.  Output goes to separate
// buffers, bucket after bucket, keys of a bucket in input order.

// Bucket counts of a multisplit in one read of the keys, counted in
// kCountCopies copies of the 256 counters, lane l using copy l % 32 (its
// bank), so the counts add no conflict across lanes.  Each workgroup first
// refines the bucket table 4x in LDS (kFineTab entries over the same image
// range, each searched only within the range its coarse entry gives), so
// that only keys in the few entries that hold a splitter search the
// splitter images; keys below the smallest splitter are bucket 0 at once.
constexpr int kSplitHistThreads = 1024;
constexpr uint32_t kCountCopies = 32;
constexpr uint32_t kFineTab = 4 * kSplitTab;

template <typename U>
__global__ __launch_bounds__(kSplitHistThreads) void split_hist_kernel(
    const U* __restrict__ keys, uint64_t n, Xform<U> xf, const SplitTab<U>* __restrict__ T, int upper,
    unsigned long long* __restrict__ g_hist, uint4* __restrict__ zero_ptr, uint64_t zero_vec4) {
  __shared__ __attribute__((aligned(16))) uint32_t s_cnt[kRadix * kCountCopies];  // [bucket][copy]
  __shared__ __attribute__((aligned(16))) uint16_t s_ent[kSplitTab];
  __shared__ __attribute__((aligned(16))) uint16_t s_fine[kFineTab];
  __shared__ U s_img[kRadix];
  const uint32_t tid = threadIdx.x;
  const uint32_t copy = tid & (kCountCopies - 1);
  uint4* c4 = reinterpret_cast<uint4*>(s_cnt);
  for (uint32_t i = tid; i < kRadix * kCountCopies / 4; i += kSplitHistThreads) c4[i] = make_uint4(0, 0, 0, 0);
  load_split_tab<U>(s_ent, s_img, T, tid, kSplitHistThreads);
  const U tab_base = T->base;
  const uint32_t tab_shift = T->shift;
  const uint32_t nspl = T->nspl;
  const uint64_t gtid = (uint64_t)blockIdx.x * kSplitHistThreads + tid;
  const uint64_t gstride = (uint64_t)gridDim.x * kSplitHistThreads;
  for (uint64_t i = gtid; i < zero_vec4; i += gstride) zero_ptr[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  // Fine entry f covers images [base + (f << fshift), base + ((f + 1) << fshift) - 1]
  // (the last reachable entry up to the largest image); its splitter counts
  // lie within those of the coarse entries holding its two ends.
  const uint32_t fshift = tab_shift >= 2 ? tab_shift - 2 : 0;
  const U flast = (U)(U(~U(0)) - tab_base) >> fshift;  // entry of the largest image, before the clamp
  for (uint32_t f = tid; f < kFineTab; f += kSplitHistThreads) {
    uint32_t lo, hi;
    if (f != 0 && (U)f > flast) {
      lo = hi = nspl;  // no image maps here
    } else {
      const U first = (U)(tab_base + ((U)f << fshift));
      const bool to_top = f == kFineTab - 1 || (U)f == flast;
      const U end = to_top ? U(~U(0)) : (U)(tab_base + ((U)(f + 1) << fshift) - 1);
      const uint32_t ef = s_ent[tab_entry_of<U>(first, tab_base, tab_shift)];
      const uint32_t ee = s_ent[tab_entry_of<U>(end, tab_base, tab_shift)];
      lo = count_splitters<U>(s_img, ef & 0xFFu, ef >> 8, first, upper);
      hi = count_splitters<U>(s_img, lo, ee >> 8, end, upper);
    }
    s_fine[f] = (uint16_t)(lo | (hi << 8));
  }
  __syncthreads();
  auto count = [&](U key) {
    const U x = to_ordered(key, xf);
    uint32_t e = 0;  // below the smallest splitter: bucket 0
    if (x >= tab_base) {
      const U q = (U)(x - tab_base) >> fshift;
      e = s_fine[q < (U)kFineTab ? (uint32_t)q : kFineTab - 1];
    }
    uint32_t lo = e & 0xFFu, hi = e >> 8;
    while (lo < hi) {  // only keys in an entry that holds splitters
      const uint32_t mid = (lo + hi) >> 1;
      const U t = s_img[mid];
      if (upper ? (t < x) : (t <= x))
        lo = mid + 1;
      else
        hi = mid;
    }
    atomicAdd(&s_cnt[lo * kCountCopies + copy], 1u);
  };
  // 16-byte loads, two in flight per thread.
  using VT = typename Vec16<U>::T;
  constexpr int VN = Vec16<U>::N;
  const uint64_t nvec = n / VN;
  const VT* kv = reinterpret_cast<const VT*>(keys);
  uint64_t vi = gtid;
  for (; vi + gstride < nvec; vi += 2 * gstride) {
    const VT v0 = kv[vi], v1 = kv[vi + gstride];
#pragma unroll
    for (int e = 0; e < VN; ++e) count(Vec16<U>::get(v0, e));
#pragma unroll
    for (int e = 0; e < VN; ++e) count(Vec16<U>::get(v1, e));
  }
  if (vi < nvec) {
    const VT v0 = kv[vi];
#pragma unroll
    for (int e = 0; e < VN; ++e) count(Vec16<U>::get(v0, e));
  }
  if (blockIdx.x == 0)
    for (uint64_t i = nvec * VN + tid; i < n; i += kSplitHistThreads) count(keys[i]);
  __syncthreads();
  if (tid < kRadix) {
    uint32_t sum = 0;
#pragma unroll 8
    for (uint32_t c = 0; c < kCountCopies; ++c) sum += s_cnt[tid * kCountCopies + ((c + tid) & (kCountCopies - 1))];
    if (sum) atomicAdd(&g_hist[tid], (unsigned long long)sum);
  }
}

struct SplitLayout {
  size_t off_ctrl, off_hist, off_base, off_tab, off_status, total;
  uint64_t tiles;
  size_t status_word;
};

static SplitLayout split_layout(int kb, int vb, uint64_t n) {
  SplitLayout L{};
  L.tiles = (n + split_tile_keys(kb, vb) - 1) / split_tile_keys(kb, vb);
  L.status_word = (n < (1ull << 30)) ? 4 : 8;
  size_t o = 0;
  L.off_ctrl = o;
  o = align_up(o + sizeof(SortCtrl), 256);
  L.off_hist = o;
  o = align_up(o + kRadix * sizeof(unsigned long long), 256);
  L.off_base = o;
  o = align_up(o + kRadix * sizeof(uint64_t), 256);
  L.off_tab = o;
  o = align_up(o + sizeof(SplitTab<uint64_t>), 256);  // the larger of the two key widths
  L.off_status = o;
  o = align_up(o + L.tiles * kRadix * L.status_word, 256);
  L.total = o;
  return L;
}

size_t multisplit_workspace_bytes(int key_bytes, int val_bytes, uint64_t n) {
  return split_layout(key_bytes, val_bytes, n).total;
}

template <typename U, int VB, typename S>
static hipError_t multisplit_t(const U* keys, const void* vals, uint64_t n, const U* d_spl, uint32_t nspl, int upper,
                               const XformParams& xp, U* out_keys, void* out_vals, unsigned char* ws,
                               const SplitLayout& L, uint64_t* d_bucket_base, hipStream_t stream,
                               LaunchTimer* timer, int policy) {
  using V = typename ValT<VB>::T;
  using Idx = typename std::conditional<sizeof(S) == 4, uint32_t, uint64_t>::type;
  constexpr bool KV = VB != 0;
  constexpr int THREADS = SplitGeometry<U, VB>::THREADS;
  constexpr int ITEMS = SplitGeometry<U, VB>::ITEMS;
  const Xform<U> xf = xform_of<U>(xp);
  SortCtrl* ctrl = reinterpret_cast<SortCtrl*>(ws + L.off_ctrl);
  unsigned long long* hist = reinterpret_cast<unsigned long long*>(ws + L.off_hist);
  uint64_t* base = d_bucket_base ? d_bucket_base : reinterpret_cast<uint64_t*>(ws + L.off_base);
  S* status = reinterpret_cast<S*>(ws + L.off_status);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(256), 0, stream, reinterpret_cast<uint32_t*>(hist),
                     (uint64_t)(kRadix * 2), 0u);
  const uint64_t status_bytes = L.tiles * kRadix * sizeof(S);
  uint64_t blocks = (n + kSplitHistThreads * 16 - 1) / (kSplitHistThreads * 16);
  if (blocks < 1) blocks = 1;
  if (blocks > 512) blocks = 512;
  SplitTab<U>* tab = reinterpret_cast<SplitTab<U>*>(ws + L.off_tab);
  hipLaunchKernelGGL(split_tab_kernel<U>, dim3(1), dim3(kSplitTabThreads), 0, stream, d_spl, nspl, upper, xf, tab);
  if (timer) timer->begin("split_hist", n * sizeof(U) + status_bytes, stream);
  hipLaunchKernelGGL(split_hist_kernel<U>, dim3((uint32_t)blocks), dim3(kSplitHistThreads), 0, stream, keys, n, xf,
                     tab, upper, hist, reinterpret_cast<uint4*>(status), status_bytes / 16);
  if (timer) timer->end(stream);
  hipLaunchKernelGGL(radix_scan_kernel, dim3(1), dim3(256), 0, stream, hist, base, ctrl, 1, n);
  OnesweepArgs<U, V, S> a{};
  a.keys0 = const_cast<U*>(keys);
  a.keys1 = out_keys;
  a.vals0 = const_cast<V*>(reinterpret_cast<const V*>(vals));
  a.vals1 = reinterpret_cast<V*>(out_vals);
  a.n = n;
  a.xf = xf;
  a.pass = 0;
  a.shift = 0;
  a.digit_base = base;
  a.status = status;
  a.ctrl = ctrl;
  a.tiles = (uint32_t)L.tiles;
  a.split_tab = tab;
  a.upper = upper;
  a.group = tile_group(stream, policy);
  if (timer) timer->begin("multisplit_pass", 2 * n * (sizeof(U) + (KV ? sizeof(V) : 0)), stream);
  hipLaunchKernelGGL((onesweep_kernel<U, V, KV, THREADS, ITEMS, S, false, false, Idx, true>), dim3((uint32_t)L.tiles),
                     dim3(THREADS), 0, stream, a);
  if (timer) timer->end(stream);
  return hipGetLastError();
}

template <typename U, int VB>
static hipError_t multisplit_s(const U* keys, const void* vals, uint64_t n, const void* d_spl, uint32_t nspl,
                               int upper, const XformParams& xp, void* out_keys, void* out_vals, unsigned char* ws,
                               const SplitLayout& L, uint64_t* d_bucket_base, hipStream_t stream,
                               LaunchTimer* timer, int policy) {
  const U* spl = reinterpret_cast<const U*>(d_spl);
  U* ok = reinterpret_cast<U*>(out_keys);
  if (L.status_word == 4)
    return multisplit_t<U, VB, uint32_t>(keys, vals, n, spl, nspl, upper, xp, ok, out_vals, ws, L, d_bucket_base,
                                         stream, timer, policy);
  return multisplit_t<U, VB, uint64_t>(keys, vals, n, spl, nspl, upper, xp, ok, out_vals, ws, L, d_bucket_base,
                                       stream, timer, policy);
}

hipError_t multisplit(int kb, int vb, const void* keys, const void* vals, uint64_t n, const void* d_splitters,
                      uint32_t nspl, int upper, const XformParams& xf, void* out_keys, void* out_vals,
                      uint64_t* d_bucket_base, void* workspace, size_t workspace_bytes, hipStream_t stream,
                      LaunchTimer* timer, int policy) {
  if (n == 0) return hipSuccess;
  if (nspl > 255) return hipErrorInvalidValue;
  const SplitLayout L = split_layout(kb, vb, n);
  if (workspace_bytes < L.total) return hipErrorInvalidValue;
  unsigned char* ws = reinterpret_cast<unsigned char*>(workspace);
  if (kb == 4) {
    const uint32_t* k = reinterpret_cast<const uint32_t*>(keys);
    if (vb == 0)
      return multisplit_s<uint32_t, 0>(k, vals, n, d_splitters, nspl, upper, xf, out_keys, out_vals, ws, L,
                                       d_bucket_base, stream, timer, policy);
    if (vb == 4)
      return multisplit_s<uint32_t, 4>(k, vals, n, d_splitters, nspl, upper, xf, out_keys, out_vals, ws, L,
                                       d_bucket_base, stream, timer, policy);
    if (vb == 8)
      return multisplit_s<uint32_t, 8>(k, vals, n, d_splitters, nspl, upper, xf, out_keys, out_vals, ws, L,
                                       d_bucket_base, stream, timer, policy);
  } else if (kb == 8) {
    const uint64_t* k = reinterpret_cast<const uint64_t*>(keys);
    if (vb == 0)
      return multisplit_s<uint64_t, 0>(k, vals, n, d_splitters, nspl, upper, xf, out_keys, out_vals, ws, L,
                                       d_bucket_base, stream, timer, policy);
    if (vb == 4)
      return multisplit_s<uint64_t, 4>(k, vals, n, d_splitters, nspl, upper, xf, out_keys, out_vals, ws, L,
                                       d_bucket_base, stream, timer, policy);
    if (vb == 8)
      return multisplit_s<uint64_t, 8>(k, vals, n, d_splitters, nspl, upper, xf, out_keys, out_vals, ws, L,
                                       d_bucket_base, stream, timer, policy);
  }
  return hipErrorInvalidValue;
}

hipError_t fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream) {
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, p, n, v);
  return hipGetLastError();
}

XformParams make_xform(int kind, bool descending) {
  XformParams x{};
  const bool wide = key_bytes(kind) == 8;
  const uint64_t msb = wide ? (1ull << 63) : (1ull << 31);
  const uint64_t all = wide ? ~0ull : 0xFFFFFFFFull;
  x.float_mask = (kind == kF32 || kind == kF64) ? all : 0;
  x.sign_flip = (kind == kU32 || kind == kU64) ? 0 : msb;
  x.desc_mask = descending ? all : 0;
  return x;
}

}  // namespace gs
