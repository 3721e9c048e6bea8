// extern "C" entry points of libmgsort_host.so (the C++ host layer above the
// libmgsort C ABI): data generation, CPU baselines, the drop-in sort drivers
// and event helpers, for callers that bind C (ctypes, the benchmarks).
// Integer returns: 0 ok, -1 unknown dtype/distribution, -3 exception (text
// in mgh_last_error()), -4 NULL argument.
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// DataGenerator(seed).ComputeDistribution<dtype>(out, n, dist).
// dtype: "int"/"i32", "long"/"i64", "float"/"f32", "double"/"f64", "u32", "u64".
int mgh_generate(const char* dtype, const char* dist, uint32_t seed, void* out, size_t n);

// CPU baselines (reference src/sort_benchmark.cu:93-108), in place.
int mgh_paradis_sort(const char* dtype, void* keys, size_t n, int threads);
int mgh_gnu_parallel_sort(const char* dtype, void* keys, size_t n, int threads);

// Drop-in drivers on host arrays (TanasicSort, GowanlockSort), in place;
// phase times afterwards via mgh_phase_seconds("sort_phase") etc.
int mgh_tanasic_sort(const char* dtype, void* keys, size_t n, int gpu, size_t partitions, uint64_t* pivot);
int mgh_gowanlock_sort(const char* dtype, void* keys, size_t n, int gpu, size_t chunk_size, size_t merge_group_size,
                       size_t num_buffers, const char* merge_backend, int threads);
// Chunk/merge-group plan: {chunk_size, merge_group_size, chunks,
// groups pre-merged, chunks pre-merged, runs of the final merge}.
int mgh_gowanlock_plan(size_t n, size_t chunk_size, size_t merge_group_size, size_t out[6]);
double mgh_phase_seconds(const char* phase);
const char* mgh_last_error(void);

// hipEvent_t helpers (for mgs_sort_timed).
void* mgh_event_create(void);
void mgh_event_destroy(void* e);
float mgh_event_elapsed_ms(void* a, void* b);

#ifdef __cplusplus
}
#endif
