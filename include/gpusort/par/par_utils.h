// ParUtils drop-in: the static helpers of the reference's
// include/par/par_utils.h:20-219 that callers of the sorts use, for a virtual
// communicator (ranks are segments of one vector, see mpi/mpi_common.h).
#ifndef GPUSORT_AMD_PAR_UTILS_H_
#define GPUSORT_AMD_PAR_UTILS_H_

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../detail/device.h"

namespace gpusort {

class ParUtils {
 public:
  // values[i] = old_values[global_indexes[i]] (par_utils.h:22-86).  The
  // indexes are positions in the concatenation of the ranks' input; the
  // gather runs on the device.  orig_sizes is accepted for compatibility.
  template <typename Int, typename Vector>
  static void ArrangeValues(MpiComm comm, const std::vector<Int>& global_indexes, const std::vector<Int>& orig_sizes,
                            Vector* values) {
    (void)comm;
    (void)orig_sizes;
    std::vector<uint64_t> idx(global_indexes.begin(), global_indexes.end());
    detail::Arrange(idx, values);
  }

  // Empty ranks take data from their neighbours (par_utils.h:103-175); the
  // vector keeps its global order, only the rank sizes change.
  template <typename Vector>
  static void KeepNoProcessHasEmptyData(MpiComm comm, Vector* arr) {
    std::vector<uint64_t> sizes = comm.Sizes();
    if (sizes.empty()) return;
    if (comm.RankOffsets(arr->size()).back() != arr->size()) throw MpiException("rank sizes do not match the vector");
    detail::Check(gs_rebalance_sizes(sizes.data(), static_cast<uint32_t>(sizes.size())));
    comm.SetResult(sizes, comm.LastSplitters(), comm.LastRounds());
  }

  // Per-rank element counts (par_utils.h:97-101).
  template <typename T>
  static void GatherNElems(MpiComm comm, T nelem, std::vector<T>* out) {
    const std::vector<uint64_t>& s = comm.Sizes();
    out->assign(s.begin(), s.end());
    if (s.empty()) out->assign(1, nelem);
  }
};

}  // namespace gpusort

#endif  // GPUSORT_AMD_PAR_UTILS_H_
