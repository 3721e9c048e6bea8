// Error handling of the host layer.  Unlike the reference's CheckCudaError,
// a no-op in release builds (src/cuda_error.cuh:5-10), every HIP and libmgsort
// status is checked and turned into an exception carrying the call site.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "mgsort.h"

namespace mgsort {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void ThrowHip(hipError_t e, const char* what, const char* file, int line) {
  throw Error(std::string(what) + ": " + hipGetErrorString(e) + " at " + file + ":" + std::to_string(line));
}
inline void ThrowMgs(int st, const char* what, const char* file, int line) {
  throw Error(std::string(what) + ": " + mgs_status_str(st) + " (hip " + std::to_string(mgs_last_hip_error()) +
              ") at " + file + ":" + std::to_string(line));
}

}  // namespace mgsort

#define HIP_CHECK(call)                                                         \
  do {                                                                          \
    const hipError_t mgs_hip_err_ = (call);                                               \
    if (mgs_hip_err_ != hipSuccess) ::mgsort::ThrowHip(mgs_hip_err_, #call, __FILE__, __LINE__);    \
  } while (0)

#define MGS_CHECK(call)                                                         \
  do {                                                                          \
    const int mgs_status_ = (call);                                                     \
    if (mgs_status_ != MGS_OK) ::mgsort::ThrowMgs(mgs_status_, #call, __FILE__, __LINE__);      \
  } while (0)
