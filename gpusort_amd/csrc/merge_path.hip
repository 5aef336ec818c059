// Merge-path 2-way merge of sorted runs, batched over many jobs per launch.
//
// Device counterpart of OmpUtils::Merge (include/omp/omp_utils.h:54-145),
// which splits both inputs with sampled lower_bound searches and runs
// std::merge per thread.  Here the output of every job is cut into tiles of
// THREADS*ITEMS elements at exact merge-path diagonals (one binary search per
// tile boundary), each tile's two input ranges are staged in LDS with
// coalesced loads, every thread merges ITEMS outputs from its own diagonal,
// and the tile is written back coalesced.  Ties are taken from the first run,
// as std::merge does, so merging runs in input order is stable.
//
// Used by gs_merge (OmpUtils::Merge's seam) and by the merge-splits of
// SampleSort's small-input bitonic network; the rank rounds' merges of
// received buckets use the single-pass k-way merge (merge_kway.hip).
#include "gs_device.h"
#include "gs_internal.h"

namespace gs {

// Number of elements of run A among the first `diag` outputs of merge(A, B).
template <typename U, typename LoadA, typename LoadB>
__device__ inline uint64_t merge_path_search(uint64_t diag, uint64_t na, uint64_t nb, LoadA a, LoadB b) {
  uint64_t lo = diag > nb ? diag - nb : 0;
  uint64_t hi = diag < na ? diag : na;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a(mid) <= b(diag - 1 - mid))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Element compared by a merge: the ordered key, and with LEX also the ordered
// payload as the second field (the lexicographic order of std::pair under
// the lifted comparator, include/common.h:156-164).
template <typename U, typename V, bool LEX>
struct MKey {
  U k;
  V v;
  __device__ bool operator<=(const MKey& o) const {
    if constexpr (LEX) return k < o.k || (k == o.k && v <= o.v);
    return k <= o.k;
  }
};

template <typename U, typename V, bool LEX>
__global__ __launch_bounds__(256) void merge_partition_kernel(const U* __restrict__ keys_a,
                                                              const U* __restrict__ keys_b,
                                                              const V* __restrict__ vals_a,
                                                              const V* __restrict__ vals_b,
                                                              const MergeJobDev* __restrict__ jobs,
                                                              const uint32_t* __restrict__ tile_job,
                                                              uint32_t ntiles, uint32_t tile_elems, Xform<U> xf,
                                                              Xform<V> vxf, uint64_t* __restrict__ a_split) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ntiles) return;
  const MergeJobDev jb = jobs[tile_job[t]];
  const uint64_t diag = (uint64_t)(t - jb.tile0) * tile_elems;
  const U* A = keys_a + jb.a_off;
  const U* B = keys_b + jb.b_off;
  const V* AV = vals_a + jb.a_off;
  const V* BV = vals_b + jb.b_off;
  using E = MKey<U, V, LEX>;
  a_split[t] = merge_path_search<E>(
      diag, jb.a_len, jb.b_len,
      [&](uint64_t i) { return E{to_ordered(A[i], xf), LEX ? to_ordered(AV[i], vxf) : V(0)}; },
      [&](uint64_t i) { return E{to_ordered(B[i], xf), LEX ? to_ordered(BV[i], vxf) : V(0)}; });
}

template <typename U, typename V, bool KV, bool LEX, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void merge_tile_kernel(const U* __restrict__ ska, const U* __restrict__ skb,
                                                             const V* __restrict__ sva, const V* __restrict__ svb,
                                                             U* __restrict__ dk, V* __restrict__ dv,
                                                             const MergeJobDev* __restrict__ jobs,
                                                             const uint32_t* __restrict__ tile_job,
                                                             const uint64_t* __restrict__ a_split, Xform<U> xf,
                                                             Xform<V> vxf) {
  using E = MKey<U, V, LEX>;
  constexpr uint32_t TILE = THREADS * ITEMS;
  __shared__ U s_k[TILE];
  __shared__ V s_v[KV ? TILE : 1];

  const uint32_t t = blockIdx.x;
  const uint32_t tid = threadIdx.x;
  const MergeJobDev jb = jobs[tile_job[t]];
  const uint64_t L = jb.a_len + jb.b_len;
  const uint64_t d0 = (uint64_t)(t - jb.tile0) * TILE;
  const uint64_t d1 = (d0 + TILE < L) ? d0 + TILE : L;
  const uint64_t a0 = a_split[t];
  const uint64_t a1 = (d1 == L) ? jb.a_len : a_split[t + 1];
  const uint64_t b0 = d0 - a0;
  const uint32_t na = (uint32_t)(a1 - a0);
  const uint32_t tot = (uint32_t)(d1 - d0);
  const uint32_t nb = tot - na;

  for (uint32_t i = tid; i < tot; i += THREADS) {
    if (i < na) {
      const uint64_t g = jb.a_off + a0 + i;
      s_k[i] = to_ordered(ska[g], xf);
      if constexpr (KV) s_v[i] = sva[g];
    } else {
      const uint64_t g = jb.b_off + b0 + (i - na);
      s_k[i] = to_ordered(skb[g], xf);
      if constexpr (KV) s_v[i] = svb[g];
    }
  }
  __syncthreads();

  const uint32_t diag = (tid * ITEMS < tot) ? tid * ITEMS : tot;
  auto elem = [&](uint32_t i) { return E{s_k[i], LEX ? to_ordered(s_v[i], vxf) : V(0)}; };
  uint32_t ia = (uint32_t)merge_path_search<E>(
      diag, na, nb, [&](uint64_t i) { return elem((uint32_t)i); }, [&](uint64_t i) { return elem(na + (uint32_t)i); });
  uint32_t ib = diag - ia;
  U ok[ITEMS];
  uint32_t osrc[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    if (diag + k < tot) {
      const bool take_a = ib >= nb || (ia < na && elem(ia) <= elem(na + ib));
      const uint32_t src = take_a ? ia : na + ib;
      ok[k] = s_k[src];
      osrc[k] = src;
      ia += take_a ? 1u : 0u;
      ib += take_a ? 0u : 1u;
    }
  }
  V ov[KV ? ITEMS : 1];
  if constexpr (KV) {
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
      if (diag + k < tot) ov[k] = s_v[osrc[k]];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    if (diag + k < tot) {
      s_k[diag + k] = ok[k];
      if constexpr (KV) s_v[diag + k] = ov[k];
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < tot; i += THREADS) {
    dk[jb.out_off + d0 + i] = from_ordered(s_k[i], xf);
    if constexpr (KV) dv[jb.out_off + d0 + i] = s_v[i];
  }
}

template <int VB>
struct MValT {
  using T = uint32_t;
};
template <>
struct MValT<8> {
  using T = uint64_t;
};

constexpr int kMergeThreads = 256;
constexpr int kMergeItems = 8;

uint32_t merge_tile_elems() { return kMergeThreads * kMergeItems; }

template <typename U, int VB, bool LEX>
static hipError_t merge_launch(const void* ska, const void* skb, const void* sva, const void* svb, void* dk,
                               void* dv, const MergeJobDev* d_jobs, const uint32_t* d_tile_job, uint32_t ntiles,
                               const XformParams& xp, const XformParams& vp, uint64_t* d_split, hipStream_t stream) {
  using V = typename MValT<VB>::T;
  constexpr bool KV = VB != 0;
  const Xform<U> xf{(U)xp.float_mask, (U)xp.sign_flip, (U)xp.desc_mask};
  const Xform<V> vxf{(V)vp.float_mask, (V)vp.sign_flip, (V)vp.desc_mask};
  hipLaunchKernelGGL((merge_partition_kernel<U, V, LEX>), dim3((ntiles + 255) / 256), dim3(256), 0, stream,
                     reinterpret_cast<const U*>(ska), reinterpret_cast<const U*>(skb),
                     reinterpret_cast<const V*>(sva), reinterpret_cast<const V*>(svb), d_jobs, d_tile_job, ntiles,
                     merge_tile_elems(), xf, vxf, d_split);
  hipLaunchKernelGGL((merge_tile_kernel<U, V, KV, LEX, kMergeThreads, kMergeItems>), dim3(ntiles),
                     dim3(kMergeThreads), 0, stream, reinterpret_cast<const U*>(ska), reinterpret_cast<const U*>(skb),
                     reinterpret_cast<const V*>(sva), reinterpret_cast<const V*>(svb), reinterpret_cast<U*>(dk),
                     reinterpret_cast<V*>(dv), d_jobs, d_tile_job, d_split, xf, vxf);
  return hipGetLastError();
}

hipError_t merge_tiles(int kb, int vb, const void* ska, const void* skb, const void* sva, const void* svb, void* dk,
                       void* dv, const MergeJobDev* d_jobs, const uint32_t* d_tile_job, uint32_t ntiles,
                       const XformParams& xf, uint64_t* d_split, hipStream_t stream, const XformParams* lex_vxf) {
  if (ntiles == 0) return hipSuccess;
  const XformParams vz{0, 0, 0};
  const bool lex = lex_vxf != nullptr && vb != 0;
  const XformParams& vp = lex ? *lex_vxf : vz;
#define GS_MERGE_CASE(UT, VBV)                                                                                    \
  return lex ? merge_launch<UT, VBV, true>(ska, skb, sva, svb, dk, dv, d_jobs, d_tile_job, ntiles, xf, vp, d_split, \
                                           stream)                                                                \
             : merge_launch<UT, VBV, false>(ska, skb, sva, svb, dk, dv, d_jobs, d_tile_job, ntiles, xf, vp,       \
                                            d_split, stream)
  if (kb == 4) {
    if (vb == 0) return merge_launch<uint32_t, 0, false>(ska, skb, sva, svb, dk, dv, d_jobs, d_tile_job, ntiles, xf, vp,
                                                         d_split, stream);
    if (vb == 4) GS_MERGE_CASE(uint32_t, 4);
    if (vb == 8) GS_MERGE_CASE(uint32_t, 8);
  } else if (kb == 8) {
    if (vb == 0) return merge_launch<uint64_t, 0, false>(ska, skb, sva, svb, dk, dv, d_jobs, d_tile_job, ntiles, xf, vp,
                                                         d_split, stream);
    if (vb == 4) GS_MERGE_CASE(uint64_t, 4);
    if (vb == 8) GS_MERGE_CASE(uint64_t, 8);
  }
#undef GS_MERGE_CASE
  return hipErrorInvalidValue;
}

// Host-side tiling of a job list: fills the device job records and the
// tile -> job table.  Returns the number of tiles.
// ------------------------------------------------------ bitonic merge-split cuts
//
// One merge-split of SampleSort's small-input bitonic network keeps a prefix
// or a suffix of merge(a, b) (MergeLists, include/helpers/samplesort_helper.h:
// 305-387).  The reference walks the merged list with loop conditions built
// from low = max of the first elements, high = min of the last ones (in the
// natural order; last/first swap for a decreasing order) and half = |a+b|/2;
// every condition holds on a prefix (or suffix) of the sorted merge, so the
// kept length is a min/max of element counts, found here with binary searches
// on a and b, in the sort order of the ordered images (which is the natural
// order, reversed for a decreasing sort):
//   increasing, keep low   L = min(max(#(< low), half), #(<= high))     prefix
//   increasing, keep high  L = max(min(n - half, #(>= low)), #(> high)) suffix
//   decreasing, keep low   L = max(min(#(>= high'), n - half), #(> low')) suffix
//   decreasing, keep high  L = max(min(half, #(<= low')), #(< high'))   prefix
// with low' = min of the last elements and high' = max of the first ones in
// image order.  The merge-path split of the kept part is returned with it.

template <typename E, typename Load>
__device__ inline uint64_t count_less(Load f, uint64_t n, const E& x) {  // #(f(i) < x)
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (x <= f(mid))
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

template <typename E, typename Load>
__device__ inline uint64_t count_le(Load f, uint64_t n, const E& x) {  // #(f(i) <= x)
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (f(mid) <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

template <typename U, typename V, bool LEX>
__global__ __launch_bounds__(64) void net_cut_kernel(const U* __restrict__ keys, const V* __restrict__ vals,
                                                     const NetCut* __restrict__ ops, uint32_t nops, Xform<U> xf,
                                                     Xform<V> vxf, int increasing, NetCutResult* __restrict__ out) {
  const uint32_t t = blockIdx.x * 64 + threadIdx.x;
  if (t >= nops) return;
  const NetCut op = ops[t];
  using E = MKey<U, V, LEX>;
  const U* A = keys + op.a_off;
  const U* B = keys + op.b_off;
  const V* AV = vals + op.a_off;
  const V* BV = vals + op.b_off;
  auto ea = [&](uint64_t i) { return E{to_ordered(A[i], xf), LEX ? to_ordered(AV[i], vxf) : V(0)}; };
  auto eb = [&](uint64_t i) { return E{to_ordered(B[i], xf), LEX ? to_ordered(BV[i], vxf) : V(0)}; };
  const uint64_t na = op.a_len, nb = op.b_len, n = na + nb, half = n / 2;
  auto cnt_less = [&](const E& x) { return count_less<E>(ea, na, x) + count_less<E>(eb, nb, x); };
  auto cnt_le = [&](const E& x) { return count_le<E>(ea, na, x) + count_le<E>(eb, nb, x); };
  auto emax = [](const E& x, const E& y) { return x <= y ? y : x; };
  auto emin = [](const E& x, const E& y) { return x <= y ? x : y; };
  const E a0 = ea(0), b0 = eb(0), al = ea(na - 1), bl = eb(nb - 1);
  uint64_t L;
  bool prefix;
  if (increasing) {
    const E low = emax(a0, b0), high = emin(al, bl);
    if (op.keep_low) {
      const uint64_t c = cnt_less(low);
      L = (c > half ? c : half);
      const uint64_t h = cnt_le(high);
      L = L < h ? L : h;
      prefix = true;
    } else {
      const uint64_t ge = n - cnt_less(low), gt = n - cnt_le(high);
      L = (n - half) < ge ? (n - half) : ge;
      L = L > gt ? L : gt;
      prefix = false;
    }
  } else {
    const E low = emin(al, bl), high = emax(a0, b0);
    if (op.keep_low) {
      const uint64_t ge = n - cnt_less(high), gt = n - cnt_le(low);
      L = ge < (n - half) ? ge : (n - half);
      L = L > gt ? L : gt;
      prefix = false;
    } else {
      const uint64_t le = cnt_le(low), lt = cnt_less(high);
      L = half < le ? half : le;
      L = L > lt ? L : lt;
      prefix = true;
    }
  }
  const uint64_t diag = prefix ? L : n - L;
  const uint64_t ai = merge_path_search<E>(diag, na, nb, ea, eb);
  out[t] = NetCutResult{L, ai, diag - ai, prefix ? 1u : 0u, 0u};
}

hipError_t net_cuts(int kb, int vb, const void* keys, const void* vals, const NetCut* d_ops, uint32_t nops,
                    const XformParams& xf, const XformParams* lex_vxf, int increasing, NetCutResult* d_out,
                    hipStream_t stream) {
  if (nops == 0) return hipSuccess;
  const dim3 grid((nops + 63) / 64), block(64);
  const bool lex = lex_vxf != nullptr && vb != 0;
  const XformParams vz{0, 0, 0};
  const XformParams& vp = lex ? *lex_vxf : vz;
#define GS_CUT(UT, VT, LEXV)                                                                                     \
  hipLaunchKernelGGL((net_cut_kernel<UT, VT, LEXV>), grid, block, 0, stream, reinterpret_cast<const UT*>(keys),  \
                     reinterpret_cast<const VT*>(vals), d_ops, nops,                                              \
                     Xform<UT>{(UT)xf.float_mask, (UT)xf.sign_flip, (UT)xf.desc_mask},                           \
                     Xform<VT>{(VT)vp.float_mask, (VT)vp.sign_flip, (VT)vp.desc_mask}, increasing, d_out);       \
  return hipGetLastError()
  if (kb == 4) {
    if (!lex) { GS_CUT(uint32_t, uint32_t, false); }
    if (vb == 4) { GS_CUT(uint32_t, uint32_t, true); }
    if (vb == 8) { GS_CUT(uint32_t, uint64_t, true); }
  } else if (kb == 8) {
    if (!lex) { GS_CUT(uint64_t, uint32_t, false); }
    if (vb == 4) { GS_CUT(uint64_t, uint32_t, true); }
    if (vb == 8) { GS_CUT(uint64_t, uint64_t, true); }
  }
#undef GS_CUT
  return hipErrorInvalidValue;
}

uint32_t merge_plan(const MergeJob* jobs, uint32_t njobs, MergeJobDev* out_jobs, std::vector<uint32_t>* tile_job) {
  const uint32_t T = merge_tile_elems();
  tile_job->clear();
  for (uint32_t j = 0; j < njobs; ++j) {
    const uint64_t L = jobs[j].a_len + jobs[j].b_len;
    out_jobs[j].a_off = jobs[j].a_off;
    out_jobs[j].a_len = jobs[j].a_len;
    out_jobs[j].b_off = jobs[j].b_off;
    out_jobs[j].b_len = jobs[j].b_len;
    out_jobs[j].out_off = jobs[j].out_off;
    out_jobs[j].tile0 = (uint32_t)tile_job->size();
    const uint64_t nt = (L + T - 1) / T;
    for (uint64_t i = 0; i < nt; ++i) tile_job->push_back(j);
  }
  return (uint32_t)tile_job->size();
}

}  // namespace gs
