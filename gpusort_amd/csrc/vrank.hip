// Virtual-rank per-rank pipeline (gs_vrank_sort): placeholder until the round
// driver lands; reports the call as unsupported.
#include <hip/hip_runtime.h>

#include "../../include/gpusort_hip.h"
#include "gs_capi_impl.h"
#include "gs_internal.h"

using namespace gs;

extern "C" gs_status gs_vrank_sort(gs_handle* h, gs_algorithm algo, gs_dtype kdt, void* d_keys, gs_dtype vdt,
                                   void* d_vals, const uint64_t* h_rank_off, uint32_t p, uint32_t k_way,
                                   uint32_t samples_per_rank, int descending, gs_vrank_report* report) {
  (void)h;
  (void)algo;
  (void)kdt;
  (void)d_keys;
  (void)vdt;
  (void)d_vals;
  (void)h_rank_off;
  (void)p;
  (void)k_way;
  (void)samples_per_rank;
  (void)descending;
  (void)report;
  return fail(GS_UNSUPPORTED, "gs_vrank_sort is not built yet");
}
