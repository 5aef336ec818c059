// CudaUtils drop-in: the static seam the reference's helpers call for the
// local sort (include/cuda/cuda_utils.h:10-28, src/cuda_utils.cu:97-243),
// forwarding to the C ABI of libgpusort_hip.so.  Header-only: no separately
// compiled instantiation matrix is needed, any key of 4 or 8 bytes with a
// std::less / std::greater ordering instantiates here.
#ifndef GPUSORT_AMD_CUDA_UTILS_H_
#define GPUSORT_AMD_CUDA_UTILS_H_

#include <cstddef>
#include <iterator>
#include <mutex>
#include <type_traits>
#include <utility>

#include "../detail/device.h"

namespace gpusort {

class CudaUtils {
 public:
  static int GetDeviceCount() {
    int n = 0;
    detail::Check(gs_get_device_count(&n));
    return n;
  }

  // Throws ThrustException("No CUDA-capable devices were detected") when no
  // device is visible, as the reference does (src/cuda_utils.cu:373-393).
  static bool IsHavingDevices() {
    if (GetDeviceCount() <= 0) detail::Check(GS_NO_DEVICE);
    return true;
  }

  // Sorts [first, last) of contiguous host memory in place.  `rank` picked
  // the device in the reference ((rank + 1) % devices); one device is used.
  template <typename StrictWeakOrdering, typename RandomAccessIterator>
  static void Sort(StrictWeakOrdering, int rank, RandomAccessIterator first, RandomAccessIterator last) {
    typedef typename std::iterator_traits<RandomAccessIterator>::value_type T;
    constexpr gs_dtype kdt = detail::DTypeOf<T>();
    static_assert(kdt != GS_NONE, "key type must be a 4- or 8-byte integer or a float / double");
    (void)rank;
    const uint64_t n = static_cast<uint64_t>(last - first);
    if (n == 0) return;
    std::lock_guard<std::mutex> lock(detail::CallMutex());
    detail::Check(gs_sort_keys_host(detail::Handle(), kdt, &*first, n, detail::DirectionOf<StrictWeakOrdering>()));
  }

  // Stable sort of keys with their values in place; when _out is not null it
  // also receives the sorted std::pair<K, V> records.
  template <typename StrictWeakOrdering, typename RandomAccessIterator1, typename RandomAccessIterator2>
  static void SortByKey(StrictWeakOrdering, int rank, RandomAccessIterator1 key_first, RandomAccessIterator1 key_last,
                        RandomAccessIterator2 val_first, void* _out) {
    typedef typename std::iterator_traits<RandomAccessIterator1>::value_type K;
    typedef typename std::iterator_traits<RandomAccessIterator2>::value_type V;
    constexpr gs_dtype kdt = detail::DTypeOf<K>();
    constexpr gs_dtype vdt = detail::DTypeOf<V>();
    static_assert(kdt != GS_NONE && vdt != GS_NONE, "keys and values must be 4- or 8-byte numbers");
    (void)rank;
    const uint64_t n = static_cast<uint64_t>(key_last - key_first);
    if (n == 0) return;
    typedef std::pair<K, V> P;
    std::lock_guard<std::mutex> lock(detail::CallMutex());
    detail::Check(gs_sort_pairs_host(detail::Handle(), kdt, &*key_first, vdt, &*val_first, n,
                                     detail::DirectionOf<StrictWeakOrdering>(), _out,
                                     static_cast<uint32_t>(sizeof(P)), static_cast<uint32_t>(offsetof(P, second))));
  }
};

}  // namespace gpusort

#endif  // GPUSORT_AMD_CUDA_UTILS_H_
