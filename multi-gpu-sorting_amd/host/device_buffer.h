// Owning device allocation (hipMalloc) used by the drivers; the sort kernels
// themselves never allocate.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <utility>

#include "hip_check.h"

namespace mgsort {

class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes) { Allocate(bytes); }
  ~DeviceBuffer() { Free(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept { *this = std::move(o); }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      Free();
      std::swap(p_, o.p_);
      std::swap(bytes_, o.bytes_);
    }
    return *this;
  }
  void Allocate(size_t bytes) {
    Free();
    if (bytes == 0) return;
    HIP_CHECK(hipMalloc(&p_, bytes));
    bytes_ = bytes;
  }
  void Free() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    bytes_ = 0;
  }
  template <typename T = void>
  T* get() const {
    return static_cast<T*>(p_);
  }
  size_t bytes() const { return bytes_; }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

class Stream {
 public:
  Stream() { HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
  ~Stream() { (void)hipStreamDestroy(s_); }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  hipStream_t get() const { return s_; }
  operator hipStream_t() const { return s_; }
  void Sync() { HIP_CHECK(hipStreamSynchronize(s_)); }

 private:
  hipStream_t s_ = nullptr;
};

class Event {
 public:
  explicit Event(bool timing = false) {
    HIP_CHECK(hipEventCreateWithFlags(&e_, timing ? hipEventDefault : hipEventDisableTiming));
  }
  ~Event() { (void)hipEventDestroy(e_); }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  hipEvent_t get() const { return e_; }
  operator hipEvent_t() const { return e_; }
  void Record(hipStream_t s) { HIP_CHECK(hipEventRecord(e_, s)); }
  void Sync() { HIP_CHECK(hipEventSynchronize(e_)); }

 private:
  hipEvent_t e_ = nullptr;
};

}  // namespace mgsort
