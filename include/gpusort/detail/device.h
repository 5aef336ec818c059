// Shared plumbing of the C++ drop-in headers: the process-wide device handle,
// element type -> gs_dtype, comparator -> direction, status -> exception, and
// the virtual-rank call every HykSort / SampleSort / HyperQuickSort overload
// ends in.  All work runs on the GPU through libgpusort_hip.so; there is no
// host sorting path.
#ifndef GPUSORT_AMD_DETAIL_DEVICE_H_
#define GPUSORT_AMD_DETAIL_DEVICE_H_

#include <stdint.h>
#include <string.h>

#include <functional>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../gpusort_hip.h"
#include "../exceptions/mpi_exception.h"
#include "../exceptions/thrust_exception.h"
#include "../mpi/mpi_common.h"

namespace gpusort {
namespace detail {

// Element types the device path sorts: the primitives of the reference's
// instantiation matrix (src/cuda_utils.cu:249-318).
template <class T>
constexpr gs_dtype DTypeOf() {
  return std::is_floating_point<T>::value
             ? (sizeof(T) == 4 ? GS_F32 : sizeof(T) == 8 ? GS_F64 : GS_NONE)
             : (std::is_integral<T>::value && !std::is_same<T, bool>::value)
                   ? (sizeof(T) == 4 ? (std::is_signed<T>::value ? GS_I32 : GS_U32)
                                     : sizeof(T) == 8 ? (std::is_signed<T>::value ? GS_I64 : GS_U64) : GS_NONE)
                   : GS_NONE;
}

// std::less / std::greater (typed or transparent) are the orderings the
// device path runs; 0 = ascending, 1 = descending.  As in the reference, whose
// GPU path is instantiated for exactly these two (src/cuda_utils.cu:249-318).
template <class C>
struct Direction : std::integral_constant<int, -1> {};
template <class T>
struct Direction<std::less<T>> : std::integral_constant<int, 0> {};
template <class T>
struct Direction<std::greater<T>> : std::integral_constant<int, 1> {};

template <class Comp>
constexpr int DirectionOf() {
  static_assert(Direction<Comp>::value >= 0,
                "the device sort orders by std::less or std::greater; other comparators are not supported");
  return Direction<Comp>::value;
}

// A comparator argument: excludes the k_way / pointer arguments of the other
// overloads from matching the StrictWeakOrdering parameter.
template <class C>
struct IsOrdering : std::integral_constant<bool, !std::is_arithmetic<C>::value && !std::is_pointer<C>::value> {};
template <class C>
using EnableIfOrdering = typename std::enable_if<IsOrdering<C>::value>::type;

enum class Thrower { kThrust, kMpi };

// Throws the reference's exception for a failed call (messages of
// src/cuda_utils.cu:373-441 and the helpers' "No data to sort.").
inline void Check(gs_status st, Thrower empty_as = Thrower::kThrust) {
  if (st == GS_OK) return;
  const char* msg = gs_last_error();
  switch (st) {
    case GS_NO_DEVICE:
      throw ThrustException("No CUDA-capable devices were detected");
    case GS_OUT_OF_MEMORY:
      throw ThrustException("Couldn't allocate device vector");
    case GS_EMPTY_INPUT:
      if (empty_as == Thrower::kMpi) throw MpiException("No data to sort.");
      throw ThrustException("No data to sort.");
    case GS_DEVICE_ERROR:
      throw ThrustException(std::string("Some other error happened during sort: ") + (msg ? msg : ""));
    default:
      throw ThrustException(msg && *msg ? msg : "Unknown error occurred");
  }
}

// One handle per process on device 0, created on first use and kept until the
// process ends (its arena is reused by every call).  Calls are serialised.
inline std::mutex& CallMutex() {
  static std::mutex mu;
  return mu;
}
inline gs_handle* Handle() {
  static gs_handle* h = nullptr;
  if (!h) Check(gs_create(&h, 0));
  return h;
}

template <class Vector>
inline void* DataOf(Vector* v) {
  return v->size() ? static_cast<void*>(&(*v)[0]) : nullptr;
}

// Runs gs_vrank_sort_host over the communicator's ranks and stores the final
// rank sizes and splitters back into it.
template <class K>
void VrankSort(gs_algorithm algo, MpiComm comm, unsigned k_way, int desc, K* keys, gs_dtype vdt, void* vals,
               uint64_t n, Thrower empty_as) {
  constexpr gs_dtype kdt = DTypeOf<K>();
  static_assert(kdt != GS_NONE, "key type must be a 4- or 8-byte integer or a float / double");
  if (n == 0) Check(GS_EMPTY_INPUT, empty_as);
  const std::vector<uint64_t> off = comm.RankOffsets(n);
  const uint32_t p = static_cast<uint32_t>(off.size() - 1);
  std::vector<uint64_t> sizes(p, 0);
  std::vector<K> spl(p);
  gs_vrank_report rep;
  memset(&rep, 0, sizeof(rep));
  rep.max_splitters = p;
  rep.splitters = spl.data();
  rep.final_rank_sizes = sizes.data();
  std::lock_guard<std::mutex> lock(CallMutex());
  Check(gs_vrank_sort_host(Handle(), algo, kdt, keys, vdt, vals, off.data(), p, k_way, comm.SamplesPerRank(), desc,
                           &rep),
        empty_as);
  std::vector<uint64_t> images(rep.n_splitters);
  for (uint32_t i = 0; i < rep.n_splitters; ++i) {
    uint64_t b = 0;
    memcpy(&b, &spl[i], sizeof(K));
    images[i] = b;
  }
  comm.SetResult(sizes, images, static_cast<int>(rep.rounds));
}

// HykSort's key-value local sort with the GPU flag at one rank: the
// reference runs CudaUtils::SortByKey (include/helpers/hyksort_helper.h:
// 247-261 -> src/cuda_utils.cu:477, thrust::sort_by_key), which is stable by
// key, and no exchange round follows a one-rank communicator.
template <class K>
void StablePairsOneRank(MpiComm comm, int desc, K* keys, gs_dtype vdt, void* vals, uint64_t n, Thrower empty_as) {
  constexpr gs_dtype kdt = DTypeOf<K>();
  static_assert(kdt != GS_NONE, "key type must be a 4- or 8-byte integer or a float / double");
  if (n == 0) Check(GS_EMPTY_INPUT, empty_as);
  {
    std::lock_guard<std::mutex> lock(CallMutex());
    Check(gs_sort_pairs_host(Handle(), kdt, keys, vdt, vals, n, desc, nullptr, 0, 0), empty_as);
  }
  comm.SetResult(std::vector<uint64_t>(1, n), std::vector<uint64_t>(), 0);
}

// dst[i] = src[idx[i]] for a trivially copyable element of 1, 2, 4 or 8
// bytes on the device; other element types are gathered on the host.
template <class Vector>
void Arrange(const std::vector<uint64_t>& idx, Vector* values) {
  typedef typename Vector::value_type T;
  if (values->size() != idx.size()) throw MpiException("value vector size differs from the key vector size");
  if (idx.empty()) return;
  Vector src(*values);
  const size_t eb = sizeof(T);
  if (std::is_trivially_copyable<T>::value && (eb == 1 || eb == 2 || eb == 4 || eb == 8)) {
    std::lock_guard<std::mutex> lock(CallMutex());
    Check(gs_gather_host(Handle(), static_cast<int>(eb), DataOf(&src), src.size(), idx.data(), DataOf(values),
                         idx.size()));
  } else {
    for (size_t i = 0; i < idx.size(); ++i) (*values)[i] = src[idx[i]];
  }
}

}  // namespace detail
}  // namespace gpusort

#endif  // GPUSORT_AMD_DETAIL_DEVICE_H_
