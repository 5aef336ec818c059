// Single-workgroup bitonic tile sort in LDS for small inputs and for the
// small-segment class of the batched sort.
//
// Serves the same seam as the radix sort (thrust::sort / thrust::sort_by_key,
// src/cuda_utils.cu:424, :477) when n is at most one tile.  A bitonic network
// is not stable by itself, so with a payload it sorts (ordered key, original
// index) pairs: equal keys keep their input order, as sort_by_key requires.
// Keys-only sorts compare the ordered key alone (equal images are equal bit
// patterns, so their order is unobservable).
#include "gs_device.h"
#include "gs_internal.h"

namespace gs {

template <typename U, typename V, bool KV, int THREADS, int CAP>
__global__ __launch_bounds__(THREADS) void block_sort_kernel(U* __restrict__ keys, V* __restrict__ vals,
                                                             const uint64_t* __restrict__ seg_off,
                                                             const uint32_t* __restrict__ seg_ids, uint32_t n_single,
                                                             Xform<U> xf) {
  __shared__ U s_k[CAP];
  __shared__ uint32_t s_i[KV ? CAP : 1];
  __shared__ V s_v[KV ? CAP : 1];

  uint64_t lo = 0;
  uint32_t m = n_single;
  if (seg_off != nullptr) {
    const uint32_t s = seg_ids ? seg_ids[blockIdx.x] : blockIdx.x;
    lo = seg_off[s];
    m = (uint32_t)(seg_off[s + 1] - lo);
  }
  if (m <= 1) return;
  uint32_t p2 = 2;
  while (p2 < m) p2 <<= 1;

  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < p2; i += THREADS) {
    if (i < m) {
      s_k[i] = to_ordered(keys[lo + i], xf);
      if constexpr (KV) s_v[i] = vals[lo + i];
    } else {
      s_k[i] = ~U(0);
    }
    if constexpr (KV) s_i[i] = i;
  }
  __syncthreads();

  for (uint32_t size = 2; size <= p2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = tid; t < (p2 >> 1); t += THREADS) {
        const uint32_t i = 2 * t - (t & (stride - 1));
        const uint32_t j = i + stride;
        const bool asc = (i & size) == 0;
        const U ki = s_k[i], kj = s_k[j];
        bool gt;  // (ki, ii) > (kj, ij)
        if constexpr (KV) {
          const uint32_t ii = s_i[i], ij = s_i[j];
          gt = ki > kj || (ki == kj && ii > ij);
          if (gt == asc) {
            s_k[i] = kj;
            s_k[j] = ki;
            s_i[i] = ij;
            s_i[j] = ii;
          }
        } else {
          gt = ki > kj;
          const bool lt = ki < kj;
          if (asc ? gt : lt) {
            s_k[i] = kj;
            s_k[j] = ki;
          }
        }
      }
      __syncthreads();
    }
  }

  for (uint32_t i = tid; i < m; i += THREADS) {
    keys[lo + i] = from_ordered(s_k[i], xf);
    if constexpr (KV) vals[lo + i] = s_v[s_i[i]];
  }
}

template <typename U, int VB>
struct BlockCap;
template <>
struct BlockCap<uint32_t, 0> {
  static constexpr int CAP = 8192;
};
template <>
struct BlockCap<uint32_t, 4> {
  static constexpr int CAP = 4096;
};
template <>
struct BlockCap<uint32_t, 8> {
  static constexpr int CAP = 2048;
};
template <>
struct BlockCap<uint64_t, 0> {
  static constexpr int CAP = 4096;
};
template <>
struct BlockCap<uint64_t, 4> {
  static constexpr int CAP = 2048;
};
template <>
struct BlockCap<uint64_t, 8> {
  static constexpr int CAP = 2048;
};

template <int VB>
struct BValT {
  using T = uint32_t;
};
template <>
struct BValT<8> {
  using T = uint64_t;
};

uint32_t block_sort_max(int kb, int vb) {
  if (kb == 4) return vb == 0 ? BlockCap<uint32_t, 0>::CAP : vb == 4 ? BlockCap<uint32_t, 4>::CAP : BlockCap<uint32_t, 8>::CAP;
  return vb == 0 ? BlockCap<uint64_t, 0>::CAP : vb == 4 ? BlockCap<uint64_t, 4>::CAP : BlockCap<uint64_t, 8>::CAP;
}

// Segments of up to kBlockSortSmall keys take 256-thread workgroups with a
// kBlockSortSmall-key tile (12-24 KB of LDS instead of up to 48 KB).
template <typename U, int VB>
static hipError_t launch_block_sort(void* keys, void* vals, const uint64_t* d_off, const uint32_t* d_ids,
                                    uint32_t nblocks, uint32_t n, const XformParams& xp, hipStream_t stream,
                                    bool small) {
  using V = typename BValT<VB>::T;
  constexpr bool KV = VB != 0;
  constexpr int CAP = BlockCap<U, VB>::CAP;
  static_assert(CAP >= (int)kBlockSortSmall, "the small tile is the smaller one");
  const Xform<U> xf{(U)xp.float_mask, (U)xp.sign_flip, (U)xp.desc_mask};
  if (small)
    hipLaunchKernelGGL((block_sort_kernel<U, V, KV, 256, kBlockSortSmall>), dim3(nblocks), dim3(256), 0, stream,
                       reinterpret_cast<U*>(keys), reinterpret_cast<V*>(vals), d_off, d_ids, n, xf);
  else
    hipLaunchKernelGGL((block_sort_kernel<U, V, KV, 512, CAP>), dim3(nblocks), dim3(512), 0, stream,
                       reinterpret_cast<U*>(keys), reinterpret_cast<V*>(vals), d_off, d_ids, n, xf);
  return hipGetLastError();
}

static hipError_t dispatch_block_sort(int kb, int vb, void* keys, void* vals, const uint64_t* d_off,
                                      const uint32_t* d_ids, uint32_t nblocks, uint32_t n, const XformParams& xf,
                                      hipStream_t stream, bool small) {
  if (kb == 4) {
    if (vb == 0) return launch_block_sort<uint32_t, 0>(keys, vals, d_off, d_ids, nblocks, n, xf, stream, small);
    if (vb == 4) return launch_block_sort<uint32_t, 4>(keys, vals, d_off, d_ids, nblocks, n, xf, stream, small);
    if (vb == 8) return launch_block_sort<uint32_t, 8>(keys, vals, d_off, d_ids, nblocks, n, xf, stream, small);
  } else if (kb == 8) {
    if (vb == 0) return launch_block_sort<uint64_t, 0>(keys, vals, d_off, d_ids, nblocks, n, xf, stream, small);
    if (vb == 4) return launch_block_sort<uint64_t, 4>(keys, vals, d_off, d_ids, nblocks, n, xf, stream, small);
    if (vb == 8) return launch_block_sort<uint64_t, 8>(keys, vals, d_off, d_ids, nblocks, n, xf, stream, small);
  }
  return hipErrorInvalidValue;
}

hipError_t block_sort(int kb, int vb, void* keys, void* vals, uint32_t n, const XformParams& xf, hipStream_t stream) {
  if (n > block_sort_max(kb, vb)) return hipErrorInvalidValue;
  if (n <= 1) return hipSuccess;
  return dispatch_block_sort(kb, vb, keys, vals, nullptr, nullptr, 1, n, xf, stream, n <= kBlockSortSmall);
}

hipError_t block_sort_segments(int kb, int vb, void* keys, void* vals, const uint64_t* d_offsets,
                               const uint32_t* d_seg_ids, uint32_t nblocks, const XformParams& xf, hipStream_t stream,
                               uint32_t max_len) {
  if (nblocks == 0) return hipSuccess;
  const bool small = max_len != 0 && max_len <= kBlockSortSmall;
  return dispatch_block_sort(kb, vb, keys, vals, d_offsets, d_seg_ids, nblocks, 0, xf, stream, small);
}

}  // namespace gs
