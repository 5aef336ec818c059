// Shared host state behind the C ABI: the handle, its HBM arena, the error
// channel and the device scope guard.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/gpusort_hip.h"
#include "gs_internal.h"

namespace gs {

// Event-pair timing of kernel launches (gs_profile_begin / gs_profile_end).
class Profiler : public LaunchTimer {
 public:
  struct Rec {
    std::string name;
    uint64_t bytes;
    hipEvent_t start, stop;
  };
  ~Profiler() override;
  void begin(const char* name, uint64_t bytes, hipStream_t stream) override;
  void end(hipStream_t stream) override;
  void reset();
  std::vector<Rec> recs;
  bool active = false;

 private:
  std::vector<hipEvent_t> pool_;
  size_t used_ = 0;
  hipEvent_t take();
};

// Grow-only device buffers, one per use.  A slot is reallocated only when a
// call needs more bytes than it holds (after the stream drains); steady-state
// calls of a given size therefore allocate nothing.
class Arena {
 public:
  enum Slot { kRadix = 0, kStageKeys, kStageVals, kStagePairs, kMerge, kSmall, kPipeKeys, kPipeVals, kPipeAux,
              kSegAux, kSegWork, kHykWork, kSlots };
  Arena() {
    for (int i = 0; i < kSlots; ++i) {
      ptr_[i] = nullptr;
      cap_[i] = 0;
    }
  }
  ~Arena();
  hipError_t get(int slot, size_t bytes, hipStream_t stream, void** out);
  // Bytes a slot allocates when it must grow to hold `bytes` (1/8 headroom,
  // whole MiB).
  static size_t grown(size_t bytes) {
    if (bytes == 0) bytes = 256;
    const size_t b = bytes + bytes / 8;
    return (b + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
  }
  size_t bytes() const;
  void release();

 private:
  void* ptr_[kSlots];
  size_t cap_[kSlots];
};

class DeviceScope {
 public:
  explicit DeviceScope(int dev);
  ~DeviceScope();

 private:
  int prev_ = -1;
  int dev_ = 0;
};

gs_status fail(gs_status st, const char* fmt, ...);
gs_status from_hip(hipError_t e, const char* what);
bool dtype_ok(gs_dtype d);
int dtype_bytes(gs_dtype d);
int val_bytes_of(gs_dtype d);
// Resolves the no-allocation entry's fixed workspace on `device` (current).
hipError_t noalloc_setup(int device);
gs_status sort_device(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals, uint64_t n,
                      int descending);
// Segmented sort from host offsets, queued on h->stream without waiting;
// `meta` must outlive the queued work; segments of 2^30+ keys are returned
// in `huge` for the caller to sort.
gs_status sort_segments_async(gs_handle* h, gs_dtype kdt, void* d_keys, gs_dtype vdt, void* d_vals,
                              const uint64_t* off, uint32_t nseg, int descending, std::vector<unsigned char>* meta,
                              std::vector<uint32_t>* huge, const SegConfig* config = nullptr);

}  // namespace gs

struct gs_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  // Tile hand-out of the handle's onesweep passes (gs::TilePolicy); the
  // server grants grouped hand-out to one slot at a time.
  int tile_policy = gs::kTilesByStream;
  // Device bytes one host-memory sort may stage (gs_set_host_sort_budget);
  // 0: the device's free memory at the call.
  uint64_t host_budget = 0;
  gs::Arena arena;
  gs::Profiler profiler;
  gs::LaunchTimer* timer() { return profiler.active ? &profiler : nullptr; }
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  // Size-class configuration of the batched sort per (key, payload) width,
  // set by the autotuner; defaults otherwise.
  std::map<int, gs::SegConfig> seg_tuned;
  gs::SegConfig seg_config(int kb, int vb) const {
    auto it = seg_tuned.find(kb * 16 + vb);
    return it == seg_tuned.end() ? gs::default_seg_config(kb, vb) : it->second;
  }
};
