// Virtual communicator: p ranks that live as consecutive segments of ONE
// host vector and are sorted on one GPU.
//
// In the reference MpiComm is MPI_Comm (include/mpi/mpi_common.h:15) and
// every rank calls HykSort with its own vector.  Here a single call sorts all
// ranks: the vector passed to HykSort / SampleSort / HyperQuickSort is the
// concatenation of the ranks' data, and the communicator holds the rank
// sizes.  After the call, Sizes() holds each rank's final size (the
// reference's per-rank vector.size() after the sort), so rank r's result is
// [Offset(r), Offset(r) + Sizes()[r]) of the vector.
//
// MpiComm is a cheap value type (copies share state), like MPI_Comm.  A
// default-constructed communicator is one rank that owns the whole vector.
#ifndef GPUSORT_AMD_MPI_COMMON_H_
#define GPUSORT_AMD_MPI_COMMON_H_

#include <stdint.h>

#include <memory>
#include <vector>

#include "../exceptions/mpi_exception.h"

namespace gpusort {

class MpiComm {
 public:
  // One rank owning the whole vector (MPI_COMM_WORLD of a 1-process job).
  MpiComm() : s_(std::make_shared<State>()) {}

  // p virtual ranks with the given initial sizes.
  static MpiComm Virtual(const std::vector<uint64_t>& sizes) {
    if (sizes.empty()) throw MpiException("a communicator needs at least one rank");
    MpiComm c;
    c.s_->sizes = sizes;
    return c;
  }

  int Size() const { return s_->sizes.empty() ? 1 : static_cast<int>(s_->sizes.size()); }
  const std::vector<uint64_t>& Sizes() const { return s_->sizes; }
  uint64_t Offset(int rank) const {
    uint64_t o = 0;
    for (int r = 0; r < rank; ++r) o += s_->sizes[r];
    return o;
  }

  // Regular samples per rank for the HykSort splitter / HyperQuickSort pivot
  // rule (the reference draws them at random).
  uint32_t SamplesPerRank() const { return s_->samples_per_rank; }
  void SetSamplesPerRank(uint32_t s) { s_->samples_per_rank = s; }

  // Splitters of the last call, per round in order, as 64-bit key images.
  const std::vector<uint64_t>& LastSplitters() const { return s_->splitters; }
  int LastRounds() const { return s_->rounds; }

  // Used by the sort bodies: offsets p+1 for a vector of n elements.
  std::vector<uint64_t> RankOffsets(uint64_t n) const {
    if (s_->sizes.empty()) return {0, n};
    std::vector<uint64_t> off(s_->sizes.size() + 1, 0);
    for (size_t r = 0; r < s_->sizes.size(); ++r) off[r + 1] = off[r] + s_->sizes[r];
    if (off.back() != n) throw MpiException("rank sizes do not add up to the vector size");
    return off;
  }
  void SetResult(const std::vector<uint64_t>& sizes, const std::vector<uint64_t>& splitters, int rounds) {
    if (!s_->sizes.empty()) s_->sizes = sizes;
    s_->splitters = splitters;
    s_->rounds = rounds;
  }

 private:
  struct State {
    std::vector<uint64_t> sizes;  // empty: one rank, sized by the call
    uint32_t samples_per_rank = 32;
    std::vector<uint64_t> splitters;
    int rounds = 0;
  };
  std::shared_ptr<State> s_;
};

}  // namespace gpusort

#endif  // GPUSORT_AMD_MPI_COMMON_H_
