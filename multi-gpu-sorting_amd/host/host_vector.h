// Pinned (page-locked) host array: the HostVector of the reference
// (src/host_vector.cuh:6-7, thrust::host_vector with a pinned allocator),
// allocated with hipHostMalloc so H2D/D2H copies run at full link speed.
// Elements are value-initialised (zero), as thrust's are.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <utility>

#include "hip_check.h"

namespace mgsort {

template <typename T>
class HostVector {
 public:
  HostVector() = default;
  explicit HostVector(size_t n, bool zero = true) { Resize(n, zero); }
  ~HostVector() { Free(); }
  HostVector(const HostVector&) = delete;
  HostVector& operator=(const HostVector&) = delete;
  HostVector(HostVector&& o) noexcept { *this = std::move(o); }
  HostVector& operator=(HostVector&& o) noexcept {
    if (this != &o) {
      Free();
      std::swap(p_, o.p_);
      std::swap(n_, o.n_);
      std::swap(pinned_, o.pinned_);
    }
    return *this;
  }

  // Falls back to ordinary (pageable) memory when no GPU runtime is present,
  // so the CPU algorithms run on machines without a device.
  void Resize(size_t n, bool zero = true) {
    Free();
    if (n == 0) return;
    if (hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p_ = static_cast<T*>(std::aligned_alloc(4096, (n * sizeof(T) + 4095) / 4096 * 4096));
      if (p_ == nullptr) throw Error("host allocation failed");
      pinned_ = false;
    }
    n_ = n;
    if (zero) std::memset(static_cast<void*>(p_), 0, n * sizeof(T));
  }
  void Free() {
    if (p_) {
      if (pinned_)
        (void)hipHostFree(p_);
      else
        std::free(p_);
    }
    p_ = nullptr;
    n_ = 0;
    pinned_ = true;
  }
  bool pinned() const { return pinned_; }
  void swap(HostVector& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    std::swap(pinned_, o.pinned_);
  }

  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  T* begin() { return p_; }
  T* end() { return p_ + n_; }
  const T* begin() const { return p_; }
  const T* end() const { return p_ + n_; }
  T& operator[](size_t i) { return p_[i]; }
  const T& operator[](size_t i) const { return p_[i]; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
  bool pinned_ = true;
};

// Mapped pinned slot the device can write directly (the reference's
// HostAllocator for the pivot result, src/host_allocator.cuh:7-31).
class MappedSlot {
 public:
  MappedSlot() {
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_), sizeof(uint64_t), hipHostMallocMapped));
    HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&dev_), host_, 0));
  }
  ~MappedSlot() { (void)hipHostFree(host_); }
  MappedSlot(const MappedSlot&) = delete;
  MappedSlot& operator=(const MappedSlot&) = delete;
  uint64_t* device() { return dev_; }
  uint64_t value() const { return *reinterpret_cast<volatile uint64_t*>(host_); }

 private:
  uint64_t* host_ = nullptr;
  uint64_t* dev_ = nullptr;
};

}  // namespace mgsort
