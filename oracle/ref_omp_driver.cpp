// ORACLE -- test infrastructure only.
//
// Exports the reference's own CPU engine, OmpUtils::MergeSort / Merge / Scan
// (include/omp/omp_utils.h:54-245 of monotone-RK/gpusort), compiled from the
// reference source tree by oracle/Makefile into oracle/_ref/libref_omp.so.
// Used to validate oracle.cpp's restatement and as the "reference" CPU
// baseline.  Key-value sorts pack std::pair and lift the comparator exactly as
// HykSortHelper::SortByKey does on its CPU branch
// (include/helpers/hyksort_helper.h:266-271, :284-291; include/common.h:156-164).
#include <stdint.h>

#include <functional>
#include <utility>
#include <vector>

#include "omp/omp_utils.h"

using gpusort::OmpUtils;

#define REF_KEY_TYPES(X) X(i32, int32_t) X(u32, uint32_t) X(f32, float) X(i64, int64_t) X(u64, uint64_t) X(f64, double)

extern "C" {

#define REF_DEF(kn, K)                                                                      \
  void ref_merge_sort_##kn(K* a, int64_t n, int desc) {                                     \
    if (desc)                                                                               \
      OmpUtils::MergeSort(std::greater<K>(), a, a + n);                                     \
    else                                                                                    \
      OmpUtils::MergeSort(std::less<K>(), a, a + n);                                        \
  }                                                                                         \
  void ref_merge_##kn(K* a, int64_t na, K* b, int64_t nb, K* out, int desc) {               \
    if (desc)                                                                               \
      OmpUtils::Merge(std::greater<K>(), a, a + na, b, b + nb, out);                        \
    else                                                                                    \
      OmpUtils::Merge(std::less<K>(), a, a + na, b, b + nb, out);                           \
  }                                                                                         \
  void ref_merge_sort_pairs_##kn##_u32(K* k, uint32_t* v, int64_t n, int desc) {            \
    typedef std::pair<K, uint32_t> P;                                                       \
    std::vector<P> items(n);                                                                \
    for (int64_t i = 0; i < n; i++) items[i] = std::make_pair(k[i], v[i]);                  \
    if (desc)                                                                               \
      OmpUtils::MergeSort(std::greater<P>(), &items[0], &items[0] + n);                     \
    else                                                                                    \
      OmpUtils::MergeSort(std::less<P>(), &items[0], &items[0] + n);                        \
    for (int64_t i = 0; i < n; i++) {                                                       \
      k[i] = items[i].first;                                                                \
      v[i] = items[i].second;                                                               \
    }                                                                                       \
  }

REF_KEY_TYPES(REF_DEF)

void ref_scan_i64(int64_t cnt, const int64_t* a, int64_t* b) { OmpUtils::Scan(cnt, a, b); }

void ref_set_threads(int n) { OmpUtils::SetNumThreads(n); }

}  // extern "C"
