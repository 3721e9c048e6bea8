// Sorts and merges of AMD's ROCm libraries (hipCUB over rocPRIM, rocThrust):
// the MI355X counterparts of the library calls that the reference's local
// benchmarks time (src/local_gpu_sort_algorithm_benchmark.cu:48-77,
// src/local_gpu_merge_algorithm_benchmark.cu:58-76).  They are comparison
// legs of the local benchmark programs only; libmgsort does not use them.
// Each returns a hipError_t value (0 on success); the Thrust legs take their
// scratch from [pool, pool + pool_bytes) and never allocate.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace mgsort {
namespace vendor {

// hipcub::DeviceRadixSort::SortKeys, out of place (in -> out); with
// temp == nullptr it only writes the scratch size to *temp_bytes.
template <typename T>
int RadixSortKeys(void* temp, size_t* temp_bytes, const T* in, T* out, size_t n, hipStream_t s);

// hipcub::DeviceMergeSort::SortKeys in place (the reference's
// mgpu::mergesort leg: a comparison merge sort).
template <typename T>
int MergeSortKeys(void* temp, size_t* temp_bytes, T* keys, size_t n, hipStream_t s);

// thrust::sort in place.
template <typename T>
int ThrustSort(T* keys, size_t n, void* pool, size_t pool_bytes, hipStream_t s);

// thrust::merge of a[0, na) and b[0, nb) into out.
template <typename T>
int ThrustMerge(const T* a, size_t na, const T* b, size_t nb, T* out, void* pool, size_t pool_bytes, hipStream_t s);

}  // namespace vendor
}  // namespace mgsort
