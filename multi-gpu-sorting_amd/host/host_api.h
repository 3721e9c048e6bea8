// extern "C" entry points of libmgsort_host.so (the C++ host layer above the
// libmgsort C ABI): data generation, CPU baselines and the drop-in sort
// drivers, for callers that bind C (ctypes, the benchmarks).
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Fills out[0, n) with DataGenerator(seed).ComputeDistribution<dtype>(dist).
// dtype: "int"/"i32", "long"/"i64", "float"/"f32", "double"/"f64", "u32",
// "u64".  Returns 0, -1 for an unknown dtype or distribution, -4 for NULL.
int mgh_generate(const char* dtype, const char* dist, uint32_t seed, void* out, size_t n);

#ifdef __cplusplus
}
#endif
