// Device memory pool of the serving layer: one hipMalloc of the HBM budget at
// start-up, carved by a first-fit free list with coalescing.  Allocation never
// calls into HIP, so admission decisions are cheap and never block on the
// device.  Thread-safe.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>

#include "hip_check.h"

namespace mgsort {

class DevicePool {
 public:
  static constexpr size_t kAlign = 256;

  explicit DevicePool(size_t bytes) : capacity_(bytes / kAlign * kAlign) {
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&base_), capacity_));
    free_[0] = capacity_;
  }
  ~DevicePool() { (void)hipFree(base_); }
  DevicePool(const DevicePool&) = delete;
  DevicePool& operator=(const DevicePool&) = delete;

  // First fit; returns nullptr when no free range is large enough.
  char* Allocate(size_t bytes) {
    bytes = (bytes + kAlign - 1) / kAlign * kAlign;
    if (bytes == 0) bytes = kAlign;
    std::lock_guard<std::mutex> lock(mu_);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_[off] = bytes;
      in_use_ += bytes;
      return base_ + off;
    }
    return nullptr;
  }

  void Free(void* p) {
    if (p == nullptr) return;
    std::lock_guard<std::mutex> lock(mu_);
    const size_t off = static_cast<size_t>(static_cast<char*>(p) - base_);
    auto u = used_.find(off);
    if (u == used_.end()) throw Error("DevicePool::Free of a pointer the pool did not hand out");
    size_t start = off, len = u->second;
    in_use_ -= len;
    used_.erase(u);
    auto next = free_.lower_bound(start);
    if (next != free_.end() && next->first == start + len) {  // merge right
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {  // merge left
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    free_[start] = len;
  }

  char* base() const { return base_; }
  size_t capacity() const { return capacity_; }
  size_t in_use() {
    std::lock_guard<std::mutex> lock(mu_);
    return in_use_;
  }
  size_t largest_free() {
    std::lock_guard<std::mutex> lock(mu_);
    size_t m = 0;
    for (const auto& f : free_) m = f.second > m ? f.second : m;
    return m;
  }

 private:
  char* base_ = nullptr;
  size_t capacity_ = 0;
  size_t in_use_ = 0;
  std::mutex mu_;
  std::map<size_t, size_t> free_;            // offset -> length
  std::unordered_map<size_t, size_t> used_;  // offset -> length
};

}  // namespace mgsort
