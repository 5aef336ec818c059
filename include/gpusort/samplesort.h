// SampleSort drop-in: the overload set of the reference's include/samplesort_common_vector.h:11-174
// over a virtual communicator (all ranks in one call, see mpi/mpi_common.h),
// run on the GPU by gs_vrank_sort (include/gpusort_hip.h).  The reference
// runs SampleSort on the host only; here it runs on the device.  Empty input
// throws MpiException("No data to sort."), as the reference's helper does.
// Comparators are std::less / std::greater; see hyksort.h for the other
// differences from the reference.
#ifndef GPUSORT_AMD_SAMPLESORT_H_
#define GPUSORT_AMD_SAMPLESORT_H_

#include <functional>
#include <vector>

#include "detail/sorts.h"

namespace gpusort {

template <typename StrictWeakOrdering, typename Vector>
detail::EnableIfOrdering<StrictWeakOrdering> SampleSort(MpiComm comm, StrictWeakOrdering comp, Vector* keys) {
  detail::SortKeys(GS_SAMPLESORT, comm, comp, 2, keys, detail::Thrower::kMpi);
}

template <typename Vector>
void SampleSort(MpiComm comm, Vector* keys) {
  SampleSort(comm, std::less<typename Vector::value_type>(), keys);
}

template <typename StrictWeakOrdering, typename VectorK, typename VectorV>
detail::EnableIfOrdering<StrictWeakOrdering> SampleSort(MpiComm comm, StrictWeakOrdering comp, VectorK* keys, VectorV* values) {
  detail::SortPairs(GS_SAMPLESORT, comm, comp, 2, keys, values, detail::Thrower::kMpi);
}

template <typename VectorK, typename VectorV>
void SampleSort(MpiComm comm, VectorK* keys, VectorV* values) {
  SampleSort(comm, std::less<typename VectorK::value_type>(), keys, values);
}

template <typename StrictWeakOrdering, typename VectorK, typename VectorV1, typename VectorV2, typename... Vectors>
detail::EnableIfOrdering<StrictWeakOrdering> SampleSort(MpiComm comm, StrictWeakOrdering comp, VectorK* keys, VectorV1* values1, VectorV2* values2,
        Vectors*... values) {
  detail::SortMulti(GS_SAMPLESORT, comm, comp, 2, keys, detail::Thrower::kMpi, values1, values2, values...);
}

template <typename VectorK, typename VectorV1, typename VectorV2, typename... Vectors>
void SampleSort(MpiComm comm, VectorK* keys, VectorV1* values1, VectorV2* values2, Vectors*... values) {
  SampleSort(comm, std::less<typename VectorK::value_type>(), keys, values1, values2, values...);
}

template <typename StrictWeakOrdering, typename VectorK, typename Int>
detail::EnableIfOrdering<StrictWeakOrdering> SampleSortWithGlobalIndexes(MpiComm comm, StrictWeakOrdering comp, const std::vector<Int>& old_sizes,
                           VectorK* keys, std::vector<Int>* g_indexes) {
  (void)old_sizes;
  detail::SortWithGlobalIndexes(GS_SAMPLESORT, comm, comp, 2, keys, g_indexes, detail::Thrower::kMpi);
}

}  // namespace gpusort

#endif  // GPUSORT_AMD_SAMPLESORT_H_
