// Rank-level failure of a drop-in call (bad communicator sizes, empty input
// for SampleSort / HyperQuickSort); mirrors the reference's
// include/exceptions/mpi_exception.h:8-22.
#ifndef GPUSORT_AMD_MPI_EXCEPTION_H_
#define GPUSORT_AMD_MPI_EXCEPTION_H_

#include <exception>
#include <string>

namespace gpusort {

class MpiException : public std::exception {
 public:
  explicit MpiException(const std::string& message) : message_(message) {}
  const char* what() const noexcept override { return message_.c_str(); }

 private:
  std::string message_;
};

}  // namespace gpusort

#endif  // GPUSORT_AMD_MPI_EXCEPTION_H_
