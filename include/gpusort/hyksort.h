// HykSort drop-in: the overload set of the reference's
// include/hyksort_common_vector.h:10-214 over a virtual communicator (all
// ranks of the communicator in one call, see mpi/mpi_common.h), run on the
// GPU by gs_vrank_sort (include/gpusort_hip.h).
//
// Differences from the reference:
//  * one call sorts every rank of the communicator; the vector is the
//    concatenation of the ranks' data and MpiComm::Sizes() holds the final
//    per-rank sizes;
//  * splitters come from MpiComm::SamplesPerRank() regular samples per rank
//    instead of rand() sampling (include/helpers/hyksort_helper.h:382-396);
//  * this library has only the device path, so the sort runs on the GPU
//    whatever is_using_gpu says; the flag selects the key-value order at one
//    rank, as the reference's two local-sort branches differ there (below);
//  * comparators are std::less / std::greater (the orderings the reference's
//    GPU path is instantiated for);
//  * no dynamic exception specifications (removed in C++17); the same
//    exception types are thrown.
// With a communicator size or k_way that is not a power of two the call runs
// SampleSort, as the reference does (include/hyksort_common_vector.inl:31-43).
#ifndef GPUSORT_AMD_HYKSORT_H_
#define GPUSORT_AMD_HYKSORT_H_

#include <functional>
#include <vector>

#include "detail/sorts.h"

namespace gpusort {

template <typename StrictWeakOrdering, typename Vector>
detail::EnableIfOrdering<StrictWeakOrdering> HykSort(MpiComm comm, StrictWeakOrdering comp, unsigned int k_way, bool is_using_gpu, Vector* keys) {
  (void)is_using_gpu;
  detail::SortKeys(GS_HYKSORT, comm, comp, k_way, keys, detail::Thrower::kThrust);
}

template <typename Vector>
void HykSort(MpiComm comm, unsigned int k_way, bool is_using_gpu, Vector* keys) {
  HykSort(comm, std::less<typename Vector::value_type>(), k_way, is_using_gpu, keys);
}

// User element type ordered by key_func(element) (a 4/8-byte integer or a
// floating-point key); elements with equal keys keep their input order.
template <typename KeyType, typename StrictWeakOrdering, typename KeyFunc, typename Vector>
detail::EnableIfOrdering<StrictWeakOrdering> HykSort(MpiComm comm, StrictWeakOrdering comp, unsigned int k_way, bool is_using_gpu, KeyFunc key_func,
             Vector* list) {
  (void)is_using_gpu;
  detail::SortByKeyFunc<KeyType>(GS_HYKSORT, comm, comp, k_way, key_func, list, detail::Thrower::kThrust);
}

template <typename KeyType, typename KeyFunc, typename Vector>
void HykSort(MpiComm comm, unsigned int k_way, bool is_using_gpu, KeyFunc key_func, Vector* list) {
  HykSort<KeyType>(comm, std::less<KeyType>(), k_way, is_using_gpu, key_func, list);
}

// Key-value order: with one rank and is_using_gpu the reference's local sort
// is CudaUtils::SortByKey, stable by key (include/helpers/hyksort_helper.h:
// 247-261); otherwise (the CPU branch, or the exchange rounds of more than
// one rank, which compare pairs through ConvertComp) elements are ordered as
// (key, value) pairs.
template <typename StrictWeakOrdering, typename VectorK, typename VectorV>
detail::EnableIfOrdering<StrictWeakOrdering> HykSort(MpiComm comm, StrictWeakOrdering comp, unsigned int k_way, bool is_using_gpu, VectorK* keys,
             VectorV* values) {
  detail::SortPairs(GS_HYKSORT, comm, comp, k_way, keys, values, detail::Thrower::kThrust, is_using_gpu);
}

template <typename VectorK, typename VectorV>
void HykSort(MpiComm comm, unsigned int k_way, bool is_using_gpu, VectorK* keys, VectorV* values) {
  HykSort(comm, std::less<typename VectorK::value_type>(), k_way, is_using_gpu, keys, values);
}

template <typename StrictWeakOrdering, typename VectorK, typename VectorV1, typename VectorV2, typename... Vectors>
detail::EnableIfOrdering<StrictWeakOrdering> HykSort(MpiComm comm, StrictWeakOrdering comp, unsigned int k_way, bool is_using_gpu, VectorK* keys,
             VectorV1* values1, VectorV2* values2, Vectors*... values) {
  (void)is_using_gpu;
  detail::SortMulti(GS_HYKSORT, comm, comp, k_way, keys, detail::Thrower::kThrust, values1, values2, values...);
}

template <typename VectorK, typename VectorV1, typename VectorV2, typename... Vectors>
void HykSort(MpiComm comm, unsigned int k_way, bool is_using_gpu, VectorK* keys, VectorV1* values1,
             VectorV2* values2, Vectors*... values) {
  HykSort(comm, std::less<typename VectorK::value_type>(), k_way, is_using_gpu, keys, values1, values2, values...);
}

// Sorts keys and returns in g_indexes the global position (in the
// concatenation of the ranks' input) each key came from
// (include/hyksort_common_vector.inl:552-565).  old_sizes is accepted for
// compatibility; the positions come from the communicator's sizes.
template <typename StrictWeakOrdering, typename VectorK, typename Int>
detail::EnableIfOrdering<StrictWeakOrdering> HykSortWithGlobalIndexes(MpiComm comm, StrictWeakOrdering comp, unsigned int k_way, bool is_using_gpu,
                              const std::vector<Int>& old_sizes, VectorK* keys, std::vector<Int>* g_indexes) {
  (void)is_using_gpu;
  (void)old_sizes;
  detail::SortWithGlobalIndexes(GS_HYKSORT, comm, comp, k_way, keys, g_indexes, detail::Thrower::kThrust);
}

}  // namespace gpusort

#endif  // GPUSORT_AMD_HYKSORT_H_
