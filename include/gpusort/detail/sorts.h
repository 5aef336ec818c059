// Bodies shared by the HykSort / SampleSort / HyperQuickSort overloads.
//
// Orderings follow the reference's lifted comparator: with one payload
// vector the elements compare as (key, payload) pairs under std::less /
// std::greater of std::pair (include/common.h:156-164); with several payload
// vectors, or a user element type sorted through a key function, the keys are
// sorted with their int64 global index as the payload and every vector is then
// gathered by index (include/hyksort_common_vector.inl:137-162, :552-565,
// include/helpers/hyksort_helper.h:150-219).
#ifndef GPUSORT_AMD_DETAIL_SORTS_H_
#define GPUSORT_AMD_DETAIL_SORTS_H_

#include <stdint.h>

#include <numeric>
#include <type_traits>
#include <vector>

#include "device.h"

namespace gpusort {
namespace detail {

template <class Comp, class VectorK>
void SortKeys(gs_algorithm algo, MpiComm comm, Comp, unsigned k_way, VectorK* keys, Thrower t) {
  typedef typename VectorK::value_type K;
  VrankSort<K>(algo, comm, k_way, DirectionOf<Comp>(), static_cast<K*>(DataOf(keys)), GS_NONE, nullptr, keys->size(),
               t);
}

template <class Comp, class VectorK, class Int>
void SortWithGlobalIndexes(gs_algorithm algo, MpiComm comm, Comp comp, unsigned k_way, VectorK* keys,
                           std::vector<Int>* g_indexes, Thrower t);

// One payload of a device type: sorted along as the pair's second member.
// stable_one_rank: HykSort's GPU branch at one rank, which is stable by key
// rather than lexicographic (include/helpers/hyksort_helper.h:247-261).
template <class Comp, class VectorK, class VectorV>
void SortPairs(gs_algorithm algo, MpiComm comm, Comp, unsigned k_way, VectorK* keys, VectorV* values, Thrower t,
               std::true_type, bool stable_one_rank = false) {
  typedef typename VectorK::value_type K;
  typedef typename VectorV::value_type V;
  if (values->size() != keys->size()) throw MpiException("value vector size differs from the key vector size");
  if (stable_one_rank && comm.Size() == 1) {
    StablePairsOneRank<K>(comm, DirectionOf<Comp>(), static_cast<K*>(DataOf(keys)), DTypeOf<V>(), DataOf(values),
                          keys->size(), t);
    return;
  }
  VrankSort<K>(algo, comm, k_way, DirectionOf<Comp>(), static_cast<K*>(DataOf(keys)), DTypeOf<V>(), DataOf(values),
               keys->size(), t);
}

// A payload of any other element type: sorted through its global index.
template <class Comp, class VectorK, class VectorV>
void SortPairs(gs_algorithm algo, MpiComm comm, Comp comp, unsigned k_way, VectorK* keys, VectorV* values, Thrower t,
               std::false_type) {
  if (values->size() != keys->size()) throw MpiException("value vector size differs from the key vector size");
  std::vector<uint64_t> idx;
  SortWithGlobalIndexes(algo, comm, comp, k_way, keys, &idx, t);
  Arrange(idx, values);
}

template <class Comp, class VectorK, class VectorV>
void DispatchPairs(gs_algorithm algo, MpiComm comm, Comp comp, unsigned k_way, VectorK* keys, VectorV* values,
                   Thrower t, std::true_type tag, bool stable_one_rank) {
  SortPairs(algo, comm, comp, k_way, keys, values, t, tag, stable_one_rank);
}
template <class Comp, class VectorK, class VectorV>
void DispatchPairs(gs_algorithm algo, MpiComm comm, Comp comp, unsigned k_way, VectorK* keys, VectorV* values,
                   Thrower t, std::false_type tag, bool) {
  SortPairs(algo, comm, comp, k_way, keys, values, t, tag);
}

template <class Comp, class VectorK, class VectorV>
void SortPairs(gs_algorithm algo, MpiComm comm, Comp comp, unsigned k_way, VectorK* keys, VectorV* values,
               Thrower t, bool stable_one_rank = false) {
  typedef typename VectorV::value_type V;
  // A payload sorted through its global index is ordered the same either way
  // (equal keys keep index order).
  DispatchPairs(algo, comm, comp, k_way, keys, values, t, std::integral_constant<bool, DTypeOf<V>() != GS_NONE>(),
                stable_one_rank);
}

template <class Comp, class VectorK, class Int>
void SortWithGlobalIndexes(gs_algorithm algo, MpiComm comm, Comp comp, unsigned k_way, VectorK* keys,
                           std::vector<Int>* g_indexes, Thrower t) {
  static_assert(std::is_integral<Int>::value && sizeof(Int) >= 4, "global indexes are 4- or 8-byte integers");
  g_indexes->resize(keys->size());
  std::iota(g_indexes->begin(), g_indexes->end(), Int(0));
  SortPairs(algo, comm, comp, k_way, keys, g_indexes, t, std::true_type());
}

inline void ArrangeAll(const std::vector<uint64_t>&) {}
template <class V0, class... Vs>
void ArrangeAll(const std::vector<uint64_t>& idx, V0* v0, Vs*... rest) {
  Arrange(idx, v0);
  ArrangeAll(idx, rest...);
}

template <class Comp, class VectorK, class... Vectors>
void SortMulti(gs_algorithm algo, MpiComm comm, Comp comp, unsigned k_way, VectorK* keys, Thrower t,
               Vectors*... values) {
  std::vector<uint64_t> idx;
  SortWithGlobalIndexes(algo, comm, comp, k_way, keys, &idx, t);
  ArrangeAll(idx, values...);
}

template <class KeyType, class Comp, class KeyFunc, class Vector>
void SortByKeyFunc(gs_algorithm algo, MpiComm comm, Comp comp, unsigned k_way, KeyFunc key_func, Vector* list,
                   Thrower t) {
  std::vector<KeyType> keys(list->size());
  for (size_t i = 0; i < keys.size(); ++i) keys[i] = key_func((*list)[i]);
  std::vector<uint64_t> idx;
  SortWithGlobalIndexes(algo, comm, comp, k_way, &keys, &idx, t);
  Arrange(idx, list);
}

}  // namespace detail
}  // namespace gpusort

#endif  // GPUSORT_AMD_DETAIL_SORTS_H_
