// Device-side failure of a drop-in call; same type and messages as the
// reference's include/exceptions/thrust_exception.h:8-22 so callers that catch
// it keep working.
#ifndef GPUSORT_AMD_THRUST_EXCEPTION_H_
#define GPUSORT_AMD_THRUST_EXCEPTION_H_

#include <exception>
#include <string>

namespace gpusort {

class ThrustException : public std::exception {
 public:
  explicit ThrustException(const std::string& message) : message_(message) {}
  const char* what() const noexcept override { return message_.c_str(); }

 private:
  std::string message_;
};

}  // namespace gpusort

#endif  // GPUSORT_AMD_THRUST_EXCEPTION_H_
