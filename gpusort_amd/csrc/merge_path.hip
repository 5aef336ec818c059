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
// Used for the k-way merge of received buckets
// (HykSortHelper::Merge, include/helpers/hyksort_helper.h:307-332) as a tree
// of 2-way levels, and for HyperQuickSort's 2-way merge
// (include/helpers/hyperquicksort_helper.h:183-202).
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
