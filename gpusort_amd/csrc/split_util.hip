// Bucket boundaries of sorted runs, value gathers and AoS pair packing.
//
// bucket_bounds is the device form of the per-rank bucket split of a sorted
// run by a splitter set:
//   HykSort / HyperQuickSort: lower_bound  (include/helpers/hyksort_helper.h:563-568,
//                                           include/helpers/hyperquicksort_helper.h:366-368)
//   SampleSort:               upper_bound  (include/helpers/samplesort_helper.h:525-537)
// One lane per (run, splitter), a binary search over the sorted run.  With a
// payload the runs are compared as (key, payload) pairs, the order the
// reference's rounds use (std::pair under the lifted comparator,
// include/common.h:156-164).
#include <type_traits>

#include "gs_device.h"
#include "gs_internal.h"

namespace gs {

template <typename U, typename V, bool LEX>
__global__ __launch_bounds__(256) void bucket_bounds_kernel(const U* __restrict__ keys, const V* __restrict__ vals,
                                                            const uint64_t* __restrict__ run_off, uint32_t nruns,
                                                            const U* __restrict__ spl, const V* __restrict__ spl_v,
                                                            const uint32_t* __restrict__ spl_first, uint32_t nspl,
                                                            int upper, Xform<U> xf, Xform<V> vxf,
                                                            uint64_t* __restrict__ bounds) {
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t per_run = (uint64_t)nspl + 2;
  if (g >= (uint64_t)nruns * per_run) return;
  const uint32_t r = (uint32_t)(g / per_run);
  const uint32_t j = (uint32_t)(g % per_run);
  const uint64_t lo0 = run_off[r], hi0 = run_off[r + 1];
  uint64_t out;
  if (j == 0) {
    out = lo0;
  } else if (j == nspl + 1) {
    out = hi0;
  } else {
    const uint64_t si = spl_first[r] + j - 1;
    const U s = to_ordered(spl[si], xf);
    const V sv = LEX ? to_ordered(spl_v[si], vxf) : V(0);
    uint64_t lo = lo0, hi = hi0;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      const U k = to_ordered(keys[mid], xf);
      bool less, equal;
      if constexpr (LEX) {
        const V v = to_ordered(vals[mid], vxf);
        less = k < s || (k == s && v < sv);
        equal = k == s && v == sv;
      } else {
        less = k < s;
        equal = k == s;
      }
      // lower_bound: first element not less than s; upper_bound: first greater.
      const bool go_right = upper ? (less || equal) : less;
      if (go_right)
        lo = mid + 1;
      else
        hi = mid;
    }
    out = lo;
  }
  bounds[g] = out;
}

template <typename U>
static hipError_t bounds_u(int vb, const void* keys, const void* vals, const uint64_t* d_run_off, uint32_t nruns,
                           const void* d_spl, const void* d_spl_v, const uint32_t* d_spl_first, uint32_t nspl,
                           int upper, const XformParams& xp, const XformParams* vp, uint64_t* d_bounds,
                           hipStream_t stream) {
  const uint64_t total = (uint64_t)nruns * (nspl + 2);
  const uint32_t blocks = (uint32_t)((total + 255) / 256);
  const Xform<U> xf{(U)xp.float_mask, (U)xp.sign_flip, (U)xp.desc_mask};
  if (vp == nullptr || vb == 0) {
    const Xform<uint32_t> vz{0, 0, 0};
    hipLaunchKernelGGL((bucket_bounds_kernel<U, uint32_t, false>), dim3(blocks), dim3(256), 0, stream,
                       reinterpret_cast<const U*>(keys), nullptr, d_run_off, nruns, reinterpret_cast<const U*>(d_spl),
                       nullptr, d_spl_first, nspl, upper, xf, vz, d_bounds);
  } else if (vb == 4) {
    const Xform<uint32_t> vxf{(uint32_t)vp->float_mask, (uint32_t)vp->sign_flip, (uint32_t)vp->desc_mask};
    hipLaunchKernelGGL((bucket_bounds_kernel<U, uint32_t, true>), dim3(blocks), dim3(256), 0, stream,
                       reinterpret_cast<const U*>(keys), reinterpret_cast<const uint32_t*>(vals), d_run_off, nruns,
                       reinterpret_cast<const U*>(d_spl), reinterpret_cast<const uint32_t*>(d_spl_v), d_spl_first,
                       nspl, upper, xf, vxf, d_bounds);
  } else {
    const Xform<uint64_t> vxf{vp->float_mask, vp->sign_flip, vp->desc_mask};
    hipLaunchKernelGGL((bucket_bounds_kernel<U, uint64_t, true>), dim3(blocks), dim3(256), 0, stream,
                       reinterpret_cast<const U*>(keys), reinterpret_cast<const uint64_t*>(vals), d_run_off, nruns,
                       reinterpret_cast<const U*>(d_spl), reinterpret_cast<const uint64_t*>(d_spl_v), d_spl_first,
                       nspl, upper, xf, vxf, d_bounds);
  }
  return hipGetLastError();
}

hipError_t bucket_bounds(int kb, const void* keys, const uint64_t* d_run_off, uint32_t nruns, const void* d_spl,
                         const uint32_t* d_spl_first, uint32_t nspl, int upper, const XformParams& xp,
                         uint64_t* d_bounds, hipStream_t stream, int vb, const void* vals, const void* d_spl_v,
                         const XformParams* lex_vxf) {
  const uint64_t total = (uint64_t)nruns * (nspl + 2);
  if (total == 0) return hipSuccess;
  if (kb == 4)
    return bounds_u<uint32_t>(vb, keys, vals, d_run_off, nruns, d_spl, d_spl_v, d_spl_first, nspl, upper, xp, lex_vxf,
                              d_bounds, stream);
  return bounds_u<uint64_t>(vb, keys, vals, d_run_off, nruns, d_spl, d_spl_v, d_spl_first, nspl, upper, xp, lex_vxf,
                            d_bounds, stream);
}

// Gathers elements idx[i] of a device array (keys or payloads) into out[i].
template <typename T>
__global__ __launch_bounds__(256) void gather_small_kernel(const T* __restrict__ src, const uint64_t* __restrict__ idx,
                                                           T* __restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

hipError_t gather_elems(int bytes, const void* src, const uint64_t* d_idx, void* out, uint32_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (n + 255) / 256;
  if (bytes == 4)
    hipLaunchKernelGGL(gather_small_kernel<uint32_t>, dim3(blocks), dim3(256), 0, stream,
                       reinterpret_cast<const uint32_t*>(src), d_idx, reinterpret_cast<uint32_t*>(out), n);
  else
    hipLaunchKernelGGL(gather_small_kernel<uint64_t>, dim3(blocks), dim3(256), 0, stream,
                       reinterpret_cast<const uint64_t*>(src), d_idx, reinterpret_cast<uint64_t*>(out), n);
  return hipGetLastError();
}

// Flag set when some element is ordered before its predecessor under the
// transform (a run that is not non-decreasing in the requested order).  Each
// lane checks one 16-byte vector and the element before it.
template <typename T>
__global__ __launch_bounds__(256) void unsorted_flag_kernel(const T* __restrict__ a, uint64_t n, Xform<T> xf,
                                                            uint32_t* __restrict__ flag) {
  constexpr uint32_t VN = 16 / sizeof(T);
  using VT = typename std::conditional<sizeof(T) == 4, uint4, ulonglong2>::type;
  const uint64_t nvec = n / VN;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  const VT* av = reinterpret_cast<const VT*>(a);
  bool bad = false;
  for (uint64_t v = (uint64_t)blockIdx.x * 256 + threadIdx.x; v < nvec; v += stride) {
    const VT x = av[v];
    T e[VN];
    if constexpr (VN == 4) {
      e[0] = x.x;
      e[1] = x.y;
      e[2] = x.z;
      e[3] = x.w;
    } else {
      e[0] = x.x;
      e[1] = x.y;
    }
    T prev = v ? to_ordered(a[v * VN - 1], xf) : T(0);
#pragma unroll
    for (uint32_t j = 0; j < VN; ++j) {
      const T cur = to_ordered(e[j], xf);
      bad |= cur < prev;
      prev = cur;
    }
  }
  // Elements past the last whole vector.
  for (uint64_t i = nvec * VN + (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    if (i > 0) bad |= to_ordered(a[i], xf) < to_ordered(a[i - 1], xf);
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(flag, 1u);
}

hipError_t unsorted_flag(int bytes, const void* a, uint64_t n, const XformParams& xp, uint32_t* d_flag,
                         hipStream_t stream) {
  if (n < 2) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (bytes == 4) {
    const Xform<uint32_t> xf{(uint32_t)xp.float_mask, (uint32_t)xp.sign_flip, (uint32_t)xp.desc_mask};
    hipLaunchKernelGGL(unsorted_flag_kernel<uint32_t>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<const uint32_t*>(a), n, xf, d_flag);
  } else {
    const Xform<uint64_t> xf{xp.float_mask, xp.sign_flip, xp.desc_mask};
    hipLaunchKernelGGL(unsorted_flag_kernel<uint64_t>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       reinterpret_cast<const uint64_t*>(a), n, xf, d_flag);
  }
  return hipGetLastError();
}

// dst[i] = src[idx[i]]: rearranges a payload vector by sorted source indexes
// (the device form of ParUtils::ArrangeValues, include/par/par_utils.h:22-86,
// on one GPU).
template <typename T>
__global__ __launch_bounds__(256) void gather_kernel(const T* __restrict__ src, const uint64_t* __restrict__ idx,
                                                     T* __restrict__ dst, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[idx[i]];
}

hipError_t gather_values(int vb, const void* src, const uint64_t* idx, void* dst, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  switch (vb) {
    case 1:
      hipLaunchKernelGGL(gather_kernel<uint8_t>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                         reinterpret_cast<const uint8_t*>(src), idx, reinterpret_cast<uint8_t*>(dst), n);
      break;
    case 2:
      hipLaunchKernelGGL(gather_kernel<uint16_t>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                         reinterpret_cast<const uint16_t*>(src), idx, reinterpret_cast<uint16_t*>(dst), n);
      break;
    case 4:
      hipLaunchKernelGGL(gather_kernel<uint32_t>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                         reinterpret_cast<const uint32_t*>(src), idx, reinterpret_cast<uint32_t*>(dst), n);
      break;
    case 8:
      hipLaunchKernelGGL(gather_kernel<uint64_t>, dim3((uint32_t)blocks), dim3(256), 0, stream,
                         reinterpret_cast<const uint64_t*>(src), idx, reinterpret_cast<uint64_t*>(dst), n);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Packs sorted keys and payloads into an array of records laid out like
// std::pair<K, V> (key at offset 0, payload at val_offset, record stride
// bytes): the _out contract of CudaUtils::SortByKey (src/cuda_utils.cu:498-502).
__global__ __launch_bounds__(256) void pack_pairs_kernel(const unsigned char* __restrict__ k, int kb,
                                                         const unsigned char* __restrict__ v, int vb,
                                                         unsigned char* __restrict__ out, uint64_t n,
                                                         uint32_t stride_bytes, uint32_t val_offset) {
  const uint64_t gstride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += gstride) {
    unsigned char* rec = out + i * stride_bytes;
    for (uint32_t b = 0; b < stride_bytes; ++b) rec[b] = 0;
    for (int b = 0; b < kb; ++b) rec[b] = k[i * kb + b];
    for (int b = 0; b < vb; ++b) rec[val_offset + b] = v[i * vb + b];
  }
}

hipError_t pack_pairs(int kb, const void* keys, int vb, const void* vals, void* out, uint64_t n, uint32_t stride_bytes,
                      uint32_t val_offset, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(pack_pairs_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream,
                     reinterpret_cast<const unsigned char*>(keys), kb, reinterpret_cast<const unsigned char*>(vals),
                     vb, reinterpret_cast<unsigned char*>(out), n, stride_bytes, val_offset);
  return hipGetLastError();
}

hipError_t copy_bytes(void* dst, const void* src, uint64_t bytes, hipStream_t stream) {
  if (bytes == 0) return hipSuccess;
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream);
}

// Index vector i -> base + i, for sorting keys together with their global
// positions (InitGlobalIndexes, include/common.h:196-203).
__global__ __launch_bounds__(256) void iota_kernel(uint64_t* __restrict__ out, uint64_t n, uint64_t base) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) out[i] = base + i;
}

hipError_t iota_u64(uint64_t* out, uint64_t n, uint64_t base, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(iota_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, out, n, base);
  return hipGetLastError();
}

}  // namespace gs
