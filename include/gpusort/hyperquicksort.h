// HyperQuickSort drop-in: the overload set of the reference's include/hyperquicksort_common_vector.h:11-174
// over a virtual communicator (all ranks in one call, see mpi/mpi_common.h),
// run on the GPU by gs_vrank_sort (include/gpusort_hip.h).  The reference
// runs HyperQuickSort on the host only; here it runs on the device.  Empty input
// throws MpiException("No data to sort."), as the reference's helper does.
// Comparators are std::less / std::greater; see hyksort.h for the other
// differences from the reference.
#ifndef GPUSORT_AMD_HYPERQUICKSORT_H_
#define GPUSORT_AMD_HYPERQUICKSORT_H_

#include <functional>
#include <vector>

#include "detail/sorts.h"

namespace gpusort {

template <typename StrictWeakOrdering, typename Vector>
detail::EnableIfOrdering<StrictWeakOrdering> HyperQuickSort(MpiComm comm, StrictWeakOrdering comp, Vector* keys) {
  detail::SortKeys(GS_HYPERQUICKSORT, comm, comp, 2, keys, detail::Thrower::kMpi);
}

template <typename Vector>
void HyperQuickSort(MpiComm comm, Vector* keys) {
  HyperQuickSort(comm, std::less<typename Vector::value_type>(), keys);
}

template <typename StrictWeakOrdering, typename VectorK, typename VectorV>
detail::EnableIfOrdering<StrictWeakOrdering> HyperQuickSort(MpiComm comm, StrictWeakOrdering comp, VectorK* keys, VectorV* values) {
  detail::SortPairs(GS_HYPERQUICKSORT, comm, comp, 2, keys, values, detail::Thrower::kMpi);
}

template <typename VectorK, typename VectorV>
void HyperQuickSort(MpiComm comm, VectorK* keys, VectorV* values) {
  HyperQuickSort(comm, std::less<typename VectorK::value_type>(), keys, values);
}

template <typename StrictWeakOrdering, typename VectorK, typename VectorV1, typename VectorV2, typename... Vectors>
detail::EnableIfOrdering<StrictWeakOrdering> HyperQuickSort(MpiComm comm, StrictWeakOrdering comp, VectorK* keys, VectorV1* values1, VectorV2* values2,
        Vectors*... values) {
  detail::SortMulti(GS_HYPERQUICKSORT, comm, comp, 2, keys, detail::Thrower::kMpi, values1, values2, values...);
}

template <typename VectorK, typename VectorV1, typename VectorV2, typename... Vectors>
void HyperQuickSort(MpiComm comm, VectorK* keys, VectorV1* values1, VectorV2* values2, Vectors*... values) {
  HyperQuickSort(comm, std::less<typename VectorK::value_type>(), keys, values1, values2, values...);
}

template <typename StrictWeakOrdering, typename VectorK, typename Int>
detail::EnableIfOrdering<StrictWeakOrdering> HyperQuickSortWithGlobalIndexes(MpiComm comm, StrictWeakOrdering comp, const std::vector<Int>& old_sizes,
                           VectorK* keys, std::vector<Int>* g_indexes) {
  (void)old_sizes;
  detail::SortWithGlobalIndexes(GS_HYPERQUICKSORT, comm, comp, 2, keys, g_indexes, detail::Thrower::kMpi);
}

}  // namespace gpusort

#endif  // GPUSORT_AMD_HYPERQUICKSORT_H_
