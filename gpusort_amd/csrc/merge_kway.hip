// Single-pass k-way merge (k <= 8) of sorted runs, batched over many
// destinations per launch.
//
// Device counterpart of HykSortHelper::Merge (include/helpers/
// hyksort_helper.h:307-332), which merges a rank's k received runs as a
// binary tree of OmpUtils::Merge calls (log2 k passes over the data).  Here
// every element is read once and written once:
//
//   samples    every kSampleStride-th element of every run is a sample; one
//              thread per sample finds its rank among all samples of its
//              destination in merged order (a binary search over the sample
//              positions of each other run).  Every kSamplesPerPart-th sample
//              is a partition boundary: its exact split of every run (the
//              number of the run's elements that precede it) is found by a
//              binary search over the run.
//   merge      one workgroup per partition stages the partition's piece of
//              every run in LDS, merges the pieces there as a tree of
//              merge-path levels (two LDS buffers), and writes the partition
//              out in merged order.
//
// Merged order: elements compare by ordered key (with LEX, by (key,
// payload) pairs, the lexicographic order of std::pair under the lifted
// comparator, include/common.h:156-164); equal elements keep the order of
// their runs in the list, then their order within the run, which is what the
// reference's tree of std::merge calls (earlier run first) produces.
//
// Partition size bound: a partition holds kSamplesPerPart samples, and the
// piece of run i in it holds c_i of them, so the piece is shorter than
// (c_i + 1) * kSampleStride; over m runs a partition has fewer than
// (kSamplesPerPart + m) * kSampleStride elements, which is the LDS tile.
#include "gs_device.h"
#include "gs_internal.h"

namespace gs {

namespace {

constexpr uint32_t kSampleStride = kKwaySampleStride;
constexpr uint32_t kSamplesPerPart = kKwaySamplesPerPart;
constexpr uint32_t kKwayCap = (kSamplesPerPart + kKwayMax) * kSampleStride;
// 256 threads x 13 outputs: an odd number of elements between consecutive
// lanes' positions spreads a wave's LDS accesses over the banks.
constexpr int kKwayThreads = 256;
constexpr int kKwayItems = 13;
static_assert(kKwayItems * kKwayThreads >= (int)kKwayCap, "tile covers the cap");
// A merge level gives each group ceil(len / kKwayItems) threads: at most
// cap / items + (groups at the first level) of them.
static_assert((kKwayCap + kKwayItems - 1) / kKwayItems + kKwayMax / 2 <= (uint32_t)kKwayThreads,
              "every group of a level has its threads");

// Element of the merged order: ordered key and, with LEX, ordered payload.
template <typename U, typename V, bool LEX>
struct KElem {
  U k;
  V v;
  __device__ bool lt(const KElem& o) const {
    if constexpr (LEX) return k < o.k || (k == o.k && v < o.v);
    return k < o.k;
  }
  __device__ bool le(const KElem& o) const {
    if constexpr (LEX) return k < o.k || (k == o.k && v <= o.v);
    return k <= o.k;
  }
};

// Does y, an element of run ry, precede x, an element of a different run
// rx, in merged order?
template <typename E>
__device__ inline bool precedes(const E& y, uint32_t ry, const E& x, uint32_t rx) {
  return ry < rx ? y.le(x) : y.lt(x);
}

// Number of t in [0, n) with pred(t) true; pred is true on a prefix.
template <typename P>
__device__ inline uint64_t count_prefix(uint64_t n, P pred) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (pred(mid))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Largest d < n with field(d) <= g (field is non-decreasing, field(0) = 0).
template <typename F>
__device__ inline uint32_t find_dest(uint32_t n, uint32_t g, F field) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (field(mid) <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Boundaries 0 (nothing taken) and nparts (every run whole) of each
// destination.
__global__ void kway_bounds_init_kernel(const KwayDest* __restrict__ dests, const KwayTotals* __restrict__ tot,
                                        uint64_t* __restrict__ bounds) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= tot->ndests) return;
  const KwayDest& D = dests[d];
  const uint32_t nparts = dests[d].nparts;
  for (uint32_t i = 0; i < kKwayMax; ++i) {
    bounds[(uint64_t)D.bnd_base * kKwayMax + i] = 0;
    bounds[((uint64_t)D.bnd_base + nparts) * kKwayMax + i] = i < D.m ? D.run_len[i] : 0;
  }
}

// Sample g (run i, sample t of its destination) of every destination:
// its element's ordered key and payload, stored compactly at index g.
template <typename U, typename V, bool LEX>
__global__ __launch_bounds__(256) void kway_gather_kernel(const U* __restrict__ keys, const V* __restrict__ vals,
                                                          const KwayDest* __restrict__ dests,
                                                          const KwayTotals* __restrict__ tot, Xform<U> xf,
                                                          Xform<V> vxf, U* __restrict__ samp_k,
                                                          V* __restrict__ samp_v) {
  const uint32_t g = blockIdx.x * 256 + threadIdx.x;
  if (g >= tot->nsamples) return;
  const uint32_t ndests = tot->ndests;
  const uint32_t d = find_dest(ndests, g, [&](uint32_t x) { return dests[x].samp_base; });
  const KwayDest& D = dests[d];
  uint64_t t = g - D.samp_base;
  uint32_t i = 0;
  while (i + 1 < D.m && t >= D.run_len[i] / kSampleStride) {
    t -= D.run_len[i] / kSampleStride;
    ++i;
  }
  const uint64_t a = D.run_off[i] + (t + 1) * kSampleStride - 1;
  samp_k[g] = to_ordered(keys[a], xf);
  if constexpr (LEX) samp_v[g] = to_ordered(vals[a], vxf);
}

// Rank of every sample among its destination's samples in merged order
// (binary searches over the compact samples of the other runs, in lockstep),
// one lane per sample of the pass.  Every kSamplesPerPart-th sample is a
// partition boundary: it records its run and position, and for every other
// run the number of that run's samples that precede it (refined to an
// element count by kway_refine_kernel).
template <typename U, typename V, bool LEX>
__global__ __launch_bounds__(256) void kway_sample_kernel(const KwayDest* __restrict__ dests,
                                                          const KwayTotals* __restrict__ tot,
                                                          const U* __restrict__ samp_k,
                                                          const V* __restrict__ samp_v,
                                                          uint64_t* __restrict__ bounds,
                                                          uint32_t* __restrict__ bound_run) {
  using E = KElem<U, V, LEX>;
  const uint32_t gs = blockIdx.x * 256 + threadIdx.x;
  if (gs >= tot->nsamples) return;
  const uint32_t d = find_dest(tot->ndests, gs, [&](uint32_t x) { return dests[x].samp_base; });
  const KwayDest& D = dests[d];
  const uint32_t m = D.m;
  uint32_t sfirst[kKwayMax + 1];  // first compact sample of every run (relative)
  uint32_t acc = 0;
#pragma unroll
  for (uint32_t r = 0; r < kKwayMax; ++r) {
    sfirst[r] = acc;
    acc += r < m ? (uint32_t)(D.run_len[r] / kSampleStride) : 0u;
  }
  sfirst[kKwayMax] = acc;
  const uint32_t g = gs - D.samp_base;  // sample within the destination
  if (g >= acc) return;
  uint32_t i = 0;
#pragma unroll
  for (uint32_t r = 1; r < kKwayMax; ++r) i += (r < m && g >= sfirst[r]) ? 1u : 0u;
  const uint32_t t = g - sfirst[i];
  const U* sk = samp_k + D.samp_base;
  const V* sv = samp_v + D.samp_base;
  const E x{sk[g], LEX ? sv[g] : V(0)};
  uint32_t lo[kKwayMax], hi[kKwayMax];
#pragma unroll
  for (uint32_t r = 0; r < kKwayMax; ++r) {
    lo[r] = sfirst[r];
    hi[r] = (r < m && r != i) ? sfirst[r + 1] : sfirst[r];
  }
  for (;;) {
    bool more = false;
    E y[kKwayMax];
    uint32_t mid[kKwayMax];
#pragma unroll
    for (uint32_t r = 0; r < kKwayMax; ++r) {
      mid[r] = (lo[r] + hi[r]) >> 1;
      if (lo[r] < hi[r]) y[r] = E{sk[mid[r]], LEX ? sv[mid[r]] : V(0)};
    }
#pragma unroll
    for (uint32_t r = 0; r < kKwayMax; ++r) {
      if (lo[r] < hi[r]) {
        if (precedes(y[r], r, x, i))
          lo[r] = mid[r] + 1;
        else
          hi[r] = mid[r];
        more |= lo[r] < hi[r];
      }
    }
    if (!more) break;
  }
  uint32_t rank = t;
#pragma unroll
  for (uint32_t r = 0; r < kKwayMax; ++r) rank += lo[r] - sfirst[r];
  if ((rank + 1) % kSamplesPerPart != 0) return;
  const uint64_t b = (uint64_t)D.bnd_base + (rank + 1) / kSamplesPerPart;  // 1 <= local b < nparts
  uint64_t* out = bounds + b * kKwayMax;
#pragma unroll
  for (uint32_t r = 0; r < kKwayMax; ++r)
    out[r] = (r == i) ? (uint64_t)(t + 1) * kSampleStride - 1 : (uint64_t)(lo[r] - sfirst[r]);
  bound_run[b] = i;
}

// Exact split of every interior boundary: its element x (run i, position j)
// is preceded by `c` samples of run r and not by the next, so the number of
// run r's elements preceding x lies in [c * stride, (c + 1) * stride - 1];
// one lane per (boundary, other run) binary-searches that window.
template <typename U, typename V, bool LEX>
__global__ __launch_bounds__(256) void kway_refine_kernel(const U* __restrict__ keys, const V* __restrict__ vals,
                                                          const KwayDest* __restrict__ dests,
                                                          const KwayTotals* __restrict__ tot, Xform<U> xf,
                                                          Xform<V> vxf, uint64_t* __restrict__ bounds,
                                                          const uint32_t* __restrict__ bound_run) {
  using E = KElem<U, V, LEX>;
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint32_t gb = (uint32_t)(t / kKwayMax), r = (uint32_t)(t % kKwayMax);
  if (gb >= tot->nbounds) return;
  const uint32_t d = find_dest(tot->ndests, gb, [&](uint32_t x) { return dests[x].bnd_base; });
  const KwayDest& D = dests[d];
  const uint32_t b = gb - D.bnd_base;
  if (b == 0 || b >= D.nparts) return;  // first and last boundaries are fixed
  const uint32_t i = bound_run[gb];
  if (r >= D.m || r == i) return;
  uint64_t* out = bounds + (uint64_t)gb * kKwayMax;
  auto elem = [&](uint32_t rr, uint64_t pos) {
    const uint64_t a = D.run_off[rr] + pos;
    return E{to_ordered(keys[a], xf), LEX ? to_ordered(vals[a], vxf) : V(0)};
  };
  const E x = elem(i, out[i]);
  const uint64_t c = out[r];
  const uint64_t len = D.run_len[r];
  const uint64_t top = (c + 1) * kSampleStride - 1;
  uint64_t lo = c * kSampleStride, hi = top < len ? top : len;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (precedes(elem(r, mid), r, x, i))
      lo = mid + 1;
    else
      hi = mid;
  }
  out[r] = lo;
}

// Everything a merge workgroup needs of its partition, in one 128-byte
// record: the source element of each piece's slot 0 (slot e of piece r reads
// src[r] + e), the slot where each piece starts (start[m..] = total), the
// output offset and the number of runs.  A partition past the pass's count
// has total 0.
struct alignas(16) KwayPart {
  uint64_t src[kKwayMax];
  uint64_t out;
  uint32_t start[kKwayMax + 1];
  uint32_t m;
  uint32_t pad[4];
};
static_assert(sizeof(KwayPart) == 128, "one 128-byte record per partition");

// One lane per partition (of the upper bound the grid is sized by): its
// destination (binary search over the destinations' first partitions), its
// boundaries and the record above.  Done here, with every partition's
// search in flight at once, rather than by one thread of each merge
// workgroup while the others wait.
__global__ __launch_bounds__(256) void kway_part_kernel(const KwayDest* __restrict__ dests,
                                                        const KwayTotals* __restrict__ tot,
                                                        const uint64_t* __restrict__ bounds,
                                                        KwayPart* __restrict__ parts, uint32_t nparts_ub) {
  const uint32_t x = blockIdx.x * 256 + threadIdx.x;
  if (x >= nparts_ub) return;
  KwayPart P;
  for (uint32_t r = 0; r < kKwayMax; ++r) {
    P.src[r] = 0;
    P.start[r] = 0;
  }
  P.start[kKwayMax] = 0;
  P.out = 0;
  P.m = 0;
  for (int i = 0; i < 4; ++i) P.pad[i] = 0;
  if (x < tot->nparts) {
    const uint32_t d = find_dest(tot->ndests, x, [&](uint32_t y) { return dests[y].part_base; });
    const KwayDest& D = dests[d];
    const uint64_t q = (uint64_t)D.bnd_base + (x - D.part_base);
    uint32_t e = 0;
    uint64_t out = D.out_off;
    for (uint32_t r = 0; r < kKwayMax; ++r) {
      const uint64_t lo = bounds[q * kKwayMax + r];
      const uint64_t hi = bounds[(q + 1) * kKwayMax + r];
      const uint32_t len = r < D.m ? (uint32_t)(hi - lo) : 0u;
      P.src[r] = D.run_off[r] + lo - e;
      P.start[r] = e;
      e += len;
      out += r < D.m ? lo : 0;
    }
    P.start[kKwayMax] = e;
    P.out = out;
    P.m = D.m;
  }
  parts[x] = P;
}

// Stages a partition's pieces in LDS in list order (coalesced within each
// piece; every thread's loads are issued before any is used).
template <typename U, typename V, bool KV, bool LEX>
__device__ __forceinline__ void kway_stage(const U* __restrict__ sk, const V* __restrict__ sv, U* s_k, V* s_v,
                                           const uint32_t (&st)[kKwayMax + 1], const uint64_t* s_src,
                                           uint32_t total, uint32_t tid, const Xform<U>& xf, const Xform<V>& vxf) {
  constexpr int VT = kKwayItems;
  U lk[VT];
  V lv[KV ? VT : 1];
#pragma unroll
  for (int u = 0; u < VT; ++u) {
    const uint32_t e = tid + (uint32_t)u * kKwayThreads;
    if (e < total) {
      uint32_t r = 0;
#pragma unroll
      for (uint32_t c = 1; c < kKwayMax; ++c) r += e >= st[c] ? 1u : 0u;  // starts: uniform values
      const uint64_t a = s_src[r] + e;
      lk[u] = sk[a];
      if constexpr (KV) lv[u] = sv[a];
    }
  }
#pragma unroll
  for (int u = 0; u < VT; ++u) {
    const uint32_t e = tid + (uint32_t)u * kKwayThreads;
    if (e < total) {
      s_k[e] = to_ordered(lk[u], xf);
      if constexpr (KV) s_v[e] = LEX ? to_ordered(lv[u], vxf) : lv[u];
    }
  }
}

// Merge tree over the staged pieces: at level L group j (pieces [j*2^L,
// (j+1)*2^L)) is the merge of its two halves, the earlier half winning
// ties.  Each group takes ceil(len / VT) threads in group order; a thread
// produces VT consecutive positions of its group (fewer at the group's end)
// with one merge-path search for its first position, then a serial merge of
// VT steps with no branch: a step past the group's end reads within the
// slack slots, and its output is not written back.  Every level's outputs
// are gathered in registers before any is written back.
//
// TIES: equal keys compare their payloads (the (key, payload) order).
// Without it equal keys go to the earlier half; with CHECK the function
// then returns whether the merged partition has equal keys out of payload
// order (block-wide), i.e. whether it differs from the (key, payload) merge:
// a key-sorted sequence whose equal keys are in payload order is the
// (key, payload) order of its elements.
// One level of the merge tree (groups of 2 * WIDTH pieces); returns, with
// CHECK on the last level, whether this thread saw equal keys out of
// payload order.  Piece starts at or past m equal the total, so group g
// spans st[g] .. st[g + 2 * WIDTH] (indices <= 8) and groups past m are
// empty.
template <typename U, typename V, bool KV, bool TIES, bool CHECK, uint32_t WIDTH>
__device__ __forceinline__ bool kway_level(U* s_k, V* s_v, const uint32_t (&st)[kKwayMax + 1], bool last,
                                           uint32_t tid) {
  constexpr int VT = kKwayItems;
  using LU = __attribute__((address_space(3))) U;
  using LV = __attribute__((address_space(3))) V;
  LU* const lk = (LU*)s_k;
  LV* const lv = (LV*)s_v;
  bool bad = false;
  U rk[VT];
  V rv[VT];  // unused (and dropped) without payloads
  uint32_t a0 = 0, b0 = 0, b1 = 0, t0 = 0, tb = 0;
  bool mine = false;
#pragma unroll
  for (uint32_t g = 0; g < kKwayMax; g += 2 * WIDTH) {
    const uint32_t nt = (st[g + 2 * WIDTH] - st[g] + VT - 1) / VT;
    const bool in = (tid >= tb) & (tid < tb + nt);
    a0 = in ? st[g] : a0;
    b0 = in ? st[g + WIDTH] : b0;
    b1 = in ? st[g + 2 * WIDTH] : b1;
    t0 = in ? tb : t0;
    mine |= in;
    tb += nt;
  }
  uint32_t o0 = 0, cnt = 0;
  if (mine) {
    const uint32_t diag = (tid - t0) * VT;
    o0 = a0 + diag;
    cnt = b1 - o0 < (uint32_t)VT ? b1 - o0 : (uint32_t)VT;
    const uint32_t na = b0 - a0, nb = b1 - b0;
    uint32_t lo = diag > nb ? diag - nb : 0;
    {
      uint32_t hi = diag < na ? diag : na;
      LU* const sa = lk + a0;
      LU* const sb = lk + (b0 + diag - 1);
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const U xa = sa[mid], xb = *(sb - mid);
        bool le;
        if constexpr (TIES) {
          le = xa < xb;
          if (xa == xb) le = lv[(sa + mid) - lk] <= lv[(sb - mid) - lk];  // payloads only for equal keys
        } else {
          le = xa <= xb;
        }
        if (le)
          lo = mid + 1;
        else
          hi = mid;
      }
    }
    // Both heads are read at every step through the sides' LDS pointers
    // (32-bit LDS addresses), so no head is kept in registers; a side's
    // pointer at or past its end reads a later slot (or a slack slot),
    // ignored through the pointer.
    LU* pa = lk + (a0 + lo);
    LU* pb = lk + (b0 + (diag - lo));
    LU* const ea = lk + b0;
    LU* const eb = lk + b1;
#pragma unroll
    for (int j = 0; j < VT; ++j) {
      const U ka = *pa, kb = *pb;
      if constexpr (TIES) {
        const V va = lv[pa - lk], vb = lv[pb - lk];
        const bool le = (ka < kb) | ((ka == kb) & (va <= vb));
        const bool take_a = (pb >= eb) | ((pa < ea) & le);
        rk[j] = take_a ? ka : kb;
        rv[j] = take_a ? va : vb;
        pa += take_a ? 1 : 0;
        pb += take_a ? 0 : 1;
      } else {
        const bool take_a = (pb >= eb) | ((pa < ea) & (ka <= kb));
        rk[j] = take_a ? ka : kb;
        // the taken element's payload, read once through its key's slot
        if constexpr (KV) rv[j] = lv[(take_a ? pa : pb) - lk];
        pa += take_a ? 1 : 0;
        pb += take_a ? 0 : 1;
      }
    }
    if constexpr (CHECK) {
      if (last) {
#pragma unroll
        for (int j = 1; j < VT; ++j)
          bad |= ((uint32_t)j < cnt) & (rk[j] == rk[j - 1]) & (rv[j] < rv[j - 1]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < VT; ++j) {
    if ((uint32_t)j < cnt) {
      s_k[o0 + j] = rk[j];
      if constexpr (KV) s_v[o0 + j] = rv[j];
    }
  }
  __syncthreads();
  if constexpr (CHECK) {
    // the pair across the boundary with the previous thread's outputs
    if (last && cnt != 0 && o0 != 0) bad |= (s_k[o0 - 1] == rk[0]) & (s_v[o0 - 1] > rv[0]);
  }
  return bad;
}

// Merge tree over the staged pieces: at level L group j (pieces [j*2^L,
// (j+1)*2^L)) is the merge of its two halves, the earlier half winning
// ties.  Each group takes ceil(len / VT) threads in group order; a thread
// produces VT consecutive positions of its group (fewer at the group's end)
// with one merge-path search for its first position, then a serial merge of
// VT steps with no branch: a step past the group's end reads within the
// slack slots, and its output is not written back.  Every level's outputs
// are gathered in registers before any is written back.
//
// TIES: equal keys compare their payloads (the (key, payload) order).
// Without it equal keys go to the earlier half; with CHECK the function
// then returns whether the merged partition has equal keys out of payload
// order (block-wide), i.e. whether it differs from the (key, payload) merge:
// a key-sorted sequence whose equal keys are in payload order is the
// (key, payload) order of its elements.
template <typename U, typename V, bool KV, bool TIES, bool CHECK>
__device__ __forceinline__ bool kway_merge_tree(U* s_k, V* s_v, const uint32_t (&st)[kKwayMax + 1], uint32_t m,
                                                uint32_t tid) {
  static_assert(kKwayMax == 8, "three levels");
  bool bad = false;
  if (m > 1) bad |= kway_level<U, V, KV, TIES, CHECK, 1>(s_k, s_v, st, m <= 2, tid);
  if (m > 2) bad |= kway_level<U, V, KV, TIES, CHECK, 2>(s_k, s_v, st, m <= 4, tid);
  if (m > 4) bad |= kway_level<U, V, KV, TIES, CHECK, 4>(s_k, s_v, st, true, tid);
  if constexpr (CHECK) return __syncthreads_or(bad) != 0;
  return false;
}

template <typename U, typename V, bool KV, bool LEX>
__global__ __launch_bounds__(kKwayThreads) void kway_merge_kernel(const U* __restrict__ sk, const V* __restrict__ sv,
                                                                  U* __restrict__ dk, V* __restrict__ dv,
                                                                  const KwayPart* __restrict__ parts, Xform<U> xf,
                                                                  Xform<V> vxf) {
  // Elements (ordered key, payload: its ordered image with LEX) of the
  // partition.  VT slack slots past the cap: a thread's steps past its
  // group's end read at most VT - 1 slots beyond it.
  __shared__ U s_k[kKwayCap + kKwayItems];
  __shared__ V s_v[KV ? kKwayCap + kKwayItems : 1];
  __shared__ KwayPart s_part;

  const uint32_t tid = threadIdx.x;
  // The partition's record: 32 lanes, one dword each.
  if (tid < sizeof(KwayPart) / 4)
    reinterpret_cast<uint32_t*>(&s_part)[tid] = reinterpret_cast<const uint32_t*>(parts + blockIdx.x)[tid];
  __syncthreads();
  const uint32_t* const s_start = s_part.start;
  const uint64_t* const s_src = s_part.src;
  const uint32_t total = s_start[kKwayMax];
  if (total == 0) return;  // past the pass's partitions (grid sized by an upper bound)
  const uint32_t m = s_part.m;
  const uint64_t out_base = s_part.out;

  uint32_t st[kKwayMax + 1];  // piece starts, uniform
#pragma unroll
  for (uint32_t c = 0; c <= kKwayMax; ++c) st[c] = __builtin_amdgcn_readfirstlane(s_start[c]);

  kway_stage<U, V, KV, LEX>(sk, sv, s_k, s_v, st, s_src, total, tid, xf, vxf);
  __syncthreads();
  if constexpr (LEX) {
    // Keys first; merged again with payloads compared on equal keys only
    // where equal keys of different pieces came out of payload order.
    if (kway_merge_tree<U, V, KV, false, true>(s_k, s_v, st, m, tid)) {
      kway_stage<U, V, KV, LEX>(sk, sv, s_k, s_v, st, s_src, total, tid, xf, vxf);
      __syncthreads();
      kway_merge_tree<U, V, KV, true, false>(s_k, s_v, st, m, tid);
    }
  } else {
    kway_merge_tree<U, V, KV, false, false>(s_k, s_v, st, m, tid);
  }
  for (uint32_t o = tid; o < total; o += kKwayThreads) {
    dk[out_base + o] = from_ordered(s_k[o], xf);
    if constexpr (KV) dv[out_base + o] = LEX ? from_ordered(s_v[o], vxf) : s_v[o];
  }
}

template <int VB>
struct KValT {
  using T = uint32_t;
};
template <>
struct KValT<8> {
  using T = uint64_t;
};

// Device scratch layout: boundaries, boundary runs, compact sample keys and
// payloads, for the pass's upper bounds.
struct KwayScratch {
  size_t bounds, bound_run, samp_k, samp_v, parts, total;
};

KwayScratch kway_scratch(const KwaySizes& B, int kb, int vb) {
  KwayScratch L{};
  size_t o = 0;
  L.bounds = o;
  o = align_up(o + (size_t)B.nbounds * kKwayMax * sizeof(uint64_t), 256);
  L.bound_run = o;
  o = align_up(o + (size_t)B.nbounds * sizeof(uint32_t), 256);
  L.samp_k = o;
  o = align_up(o + (size_t)B.nsamples * kb, 256);
  L.samp_v = o;
  o = align_up(o + (size_t)B.nsamples * (vb ? vb : 4), 256);
  L.parts = o;
  o = align_up(o + (size_t)B.nparts * sizeof(KwayPart), 256);
  L.total = o;
  return L;
}

template <typename U, int VB, bool LEX>
hipError_t kway_launch(const void* sk, const void* sv, void* dk, void* dv, const KwaySizes& B, const KwayDevice& dev,
                       const XformParams& xp, const XformParams& vp, hipStream_t stream, LaunchTimer* timer,
                       uint64_t elems) {
  using V = typename KValT<VB>::T;
  constexpr bool KV = VB != 0;
  const Xform<U> xf{(U)xp.float_mask, (U)xp.sign_flip, (U)xp.desc_mask};
  const Xform<V> vxf{(V)vp.float_mask, (V)vp.sign_flip, (V)vp.desc_mask};
  const KwayScratch L = kway_scratch(B, sizeof(U), VB);
  char* ws = reinterpret_cast<char*>(dev.scratch);
  uint64_t* bounds = reinterpret_cast<uint64_t*>(ws + L.bounds);
  uint32_t* bound_run = reinterpret_cast<uint32_t*>(ws + L.bound_run);
  U* samp_k = reinterpret_cast<U*>(ws + L.samp_k);
  V* samp_v = reinterpret_cast<V*>(ws + L.samp_v);
  KwayPart* parts = reinterpret_cast<KwayPart*>(ws + L.parts);
  if (timer) timer->begin("kway_setup", 0, stream);
  hipLaunchKernelGGL(kway_bounds_init_kernel, dim3((B.ndests + 255) / 256), dim3(256), 0, stream, dev.dests,
                     dev.totals, bounds);
  if (B.nsamples) {
    hipLaunchKernelGGL((kway_gather_kernel<U, V, LEX>), dim3((B.nsamples + 255) / 256), dim3(256), 0, stream,
                       reinterpret_cast<const U*>(sk), reinterpret_cast<const V*>(sv), dev.dests, dev.totals, xf, vxf,
                       samp_k, samp_v);
    hipLaunchKernelGGL((kway_sample_kernel<U, V, LEX>), dim3((B.nsamples + 255) / 256), dim3(256), 0, stream,
                       dev.dests, dev.totals, samp_k, samp_v, bounds, bound_run);
    hipLaunchKernelGGL((kway_refine_kernel<U, V, LEX>), dim3(((uint64_t)B.nbounds * kKwayMax + 255) / 256), dim3(256),
                       0, stream,
                       reinterpret_cast<const U*>(sk), reinterpret_cast<const V*>(sv), dev.dests, dev.totals, xf, vxf,
                       bounds, bound_run);
  }
  hipLaunchKernelGGL(kway_part_kernel, dim3((B.nparts + 255) / 256), dim3(256), 0, stream, dev.dests, dev.totals,
                     bounds, parts, B.nparts);
  if (timer) {
    timer->end(stream);
    timer->begin("kway_merge_kernel", 2 * elems * (sizeof(U) + VB), stream);
  }
  hipLaunchKernelGGL((kway_merge_kernel<U, V, KV, LEX>), dim3(B.nparts), dim3(kKwayThreads), 0, stream,
                     reinterpret_cast<const U*>(sk), reinterpret_cast<const V*>(sv), reinterpret_cast<U*>(dk),
                     reinterpret_cast<V*>(dv), parts, xf, vxf);
  if (timer) timer->end(stream);
  return hipGetLastError();
}

}  // namespace

void kway_plan(KwayPlan* P) {
  uint32_t samp = 0, bnd = 0, part = 0;
  for (size_t d = 0; d < P->dests.size(); ++d) {
    KwayDest& D = P->dests[d];
    uint64_t ns = 0;
    for (uint32_t r = 0; r < D.m; ++r) ns += D.run_len[r] / kSampleStride;
    for (uint32_t r = D.m; r < kKwayMax; ++r) {
      D.run_off[r] = 0;
      D.run_len[r] = 0;
    }
    D.samp_base = samp;
    D.bnd_base = bnd;
    D.part_base = part;
    D.nparts = (uint32_t)(ns / kSamplesPerPart) + 1;
    samp += (uint32_t)ns;
    bnd += D.nparts + 1;
    part += D.nparts;
  }
  P->totals.ndests = (uint32_t)P->dests.size();
  P->totals.nsamples = samp;
  P->totals.nbounds = bnd;
  P->totals.nparts = part;
}

KwaySizes kway_upper_bounds(uint64_t n_elems, uint32_t ndests) {
  KwaySizes B;
  B.ndests = ndests;
  B.nsamples = (uint32_t)(n_elems / kSampleStride);
  B.nparts = (uint32_t)(n_elems / ((uint64_t)kSampleStride * kSamplesPerPart)) + ndests;
  B.nbounds = B.nparts + ndests;
  return B;
}

size_t kway_scratch_bytes(const KwaySizes& B, int kb, int vb) { return kway_scratch(B, kb, vb).total; }

hipError_t kway_merge(int kb, int vb, const void* sk, const void* sv, void* dk, void* dv, const KwaySizes& B,
                      const KwayDevice& dev, const XformParams& xf, const XformParams* lex_vxf, hipStream_t stream,
                      LaunchTimer* timer, uint64_t elems) {
  if (B.ndests == 0 || B.nparts == 0) return hipSuccess;
  const XformParams vz{0, 0, 0};
  const bool lex = lex_vxf != nullptr && vb != 0;
  const XformParams& vp = lex ? *lex_vxf : vz;
#define GS_KWAY_CASE(UT, VBV)                                                                  \
  return lex ? kway_launch<UT, VBV, true>(sk, sv, dk, dv, B, dev, xf, vp, stream, timer, elems)              \
             : kway_launch<UT, VBV, false>(sk, sv, dk, dv, B, dev, xf, vp, stream, timer, elems)
  if (kb == 4) {
    if (vb == 0) return kway_launch<uint32_t, 0, false>(sk, sv, dk, dv, B, dev, xf, vp, stream, timer, elems);
    if (vb == 4) GS_KWAY_CASE(uint32_t, 4);
    if (vb == 8) GS_KWAY_CASE(uint32_t, 8);
  } else if (kb == 8) {
    if (vb == 0) return kway_launch<uint64_t, 0, false>(sk, sv, dk, dv, B, dev, xf, vp, stream, timer, elems);
    if (vb == 4) GS_KWAY_CASE(uint64_t, 4);
    if (vb == 8) GS_KWAY_CASE(uint64_t, 8);
  }
#undef GS_KWAY_CASE
  return hipErrorInvalidValue;
}

}  // namespace gs
