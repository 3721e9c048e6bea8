// The library legs of the local benchmark programs (declared in
// vendor_sorts.h), instantiated for 32- and 64-bit signed keys.
#include "vendor_sorts.h"

#include <hipcub/device/device_merge_sort.hpp>
#include <hipcub/device/device_radix_sort.hpp>
#include <thrust/execution_policy.h>
#include <thrust/merge.h>
#include <thrust/sort.h>

#include <cstdint>
#include <functional>
#include <new>

namespace mgsort {
namespace vendor {
namespace {

// Bump allocator over a buffer the caller owns (the role of the
// reference's DeviceAllocator, src/device_allocator.cuh:59-81, handed to
// thrust as the execution policy's allocator): allocations are 256-byte
// aligned, deallocation is a no-op, and the pool is reset per call.
struct PoolAllocator {
  using value_type = char;
  char* base;
  size_t cap;
  size_t used;
  char* allocate(std::ptrdiff_t n) {
    const size_t at = (used + 255) & ~size_t(255);
    if (n < 0 || at + (size_t)n > cap) throw std::bad_alloc();
    used = at + (size_t)n;
    return base + at;
  }
  void deallocate(char*, size_t) {}
};

template <typename F>
int Guarded(F&& f) {
  try {
    f();
  } catch (const std::bad_alloc&) {
    return (int)hipErrorOutOfMemory;
  } catch (...) {
    return (int)hipErrorUnknown;
  }
  return (int)hipPeekAtLastError();
}

}  // namespace

template <typename T>
int RadixSortKeys(void* temp, size_t* temp_bytes, const T* in, T* out, size_t n, hipStream_t s) {
  return (int)hipcub::DeviceRadixSort::SortKeys(temp, *temp_bytes, in, out, n, 0, (int)sizeof(T) * 8, s);
}

template <typename T>
int MergeSortKeys(void* temp, size_t* temp_bytes, T* keys, size_t n, hipStream_t s) {
  return (int)hipcub::DeviceMergeSort::SortKeys(temp, *temp_bytes, keys, n, std::less<T>(), s);
}

template <typename T>
int ThrustSort(T* keys, size_t n, void* pool, size_t pool_bytes, hipStream_t s) {
  PoolAllocator alloc{static_cast<char*>(pool), pool_bytes, 0};
  return Guarded([&] { thrust::sort(thrust::hip::par(alloc).on(s), keys, keys + n); });
}

template <typename T>
int ThrustMerge(const T* a, size_t na, const T* b, size_t nb, T* out, void* pool, size_t pool_bytes, hipStream_t s) {
  PoolAllocator alloc{static_cast<char*>(pool), pool_bytes, 0};
  return Guarded([&] { thrust::merge(thrust::hip::par(alloc).on(s), a, a + na, b, b + nb, out); });
}

#define MGS_VENDOR_INSTANTIATE(T)                                                                    \
  template int RadixSortKeys<T>(void*, size_t*, const T*, T*, size_t, hipStream_t);                 \
  template int MergeSortKeys<T>(void*, size_t*, T*, size_t, hipStream_t);                            \
  template int ThrustSort<T>(T*, size_t, void*, size_t, hipStream_t);                                \
  template int ThrustMerge<T>(const T*, size_t, const T*, size_t, T*, void*, size_t, hipStream_t);
MGS_VENDOR_INSTANTIATE(int32_t)
MGS_VENDOR_INSTANTIATE(int64_t)
#undef MGS_VENDOR_INSTANTIATE

}  // namespace vendor
}  // namespace mgsort
