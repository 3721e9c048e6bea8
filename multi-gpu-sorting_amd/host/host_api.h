// extern "C" entry points of libmgsort_host.so (the C++ host layer above the
// libmgsort C ABI): data generation, CPU baselines, the drop-in sort drivers
// and event helpers, for callers that bind C (ctypes, the benchmarks).
// Integer returns: 0 ok, -1 unknown dtype/distribution, -3 exception (text
// in mgh_last_error()), -4 NULL argument.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "mgsort.h"

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
// block_bytes: in-place transfer block of the 3-buffer schedule (0 = a
// hundredth of a chunk).
int mgh_gowanlock_sort(const char* dtype, void* keys, size_t n, int gpu, size_t chunk_size, size_t merge_group_size,
                       size_t num_buffers, const char* merge_backend, int threads, size_t block_bytes);
// Chunk/merge-group plan: {chunk_size, merge_group_size, chunks,
// groups pre-merged, chunks pre-merged, runs of the final merge}.
int mgh_gowanlock_plan(size_t n, size_t chunk_size, size_t merge_group_size, size_t out[6]);
double mgh_phase_seconds(const char* phase);
const char* mgh_last_error(void);

// One request of mgh_serve (the serving layer's SortServer, no reference
// counterpart).  Keys and payloads in host memory; the output arrays may be
// the input arrays (in place).
typedef struct mgh_serve_request {
  int key_type;         // mgs_key
  int val_type;         // mgs_val
  size_t n;
  const void* keys_in;
  const void* vals_in;  // NULL unless val_type != MGS_VAL_NONE
  void* keys_out;
  void* vals_out;
  double arrival_s;     // submit time, seconds after the first request
  double latency_s;     // out: submit -> completion event
  int status;           // out: mgs status of the request's batch
} mgh_serve_request;

// Runs `count` requests through one SortServer on device `gpu`: batched
// (continuous batching into segmented launches) or plain (one mgs_sort per
// request); host_data 1: the server copies host arrays in and out; 0: the
// requests' data is placed in the server's HBM pool first (as a client that
// already holds device data) and read back after the run.  *gkeys_s gets the
// throughput over the run's wall time.  Returns 0 when every request
// completed with MGS_OK.
int mgh_serve(mgh_serve_request* reqs, size_t count, int gpu, int batched, int max_inflight, int host_data,
              double budget_gib, double* gkeys_s);

// Host link probe on device `gpu` (no reference code is run: the modes mirror
// src/data_transfer_benchmark.cu:83-173 and the in-place transfer of
// src/inplace_memcpy.cuh:7-69).  bytes of pinned host memory and of device
// memory, `reps` timed repetitions (best reported).  mode 0: host to device;
// 1: device to host; 2: both at once on two streams (GB/s counts both
// directions); 3: in place -- bytes leave a device buffer to the host while
// other bytes refill it, block by block on two streams (block_bytes; GB/s
// counts both directions).  *gbps gets the rate.
int mgh_link_probe(int gpu, int mode, size_t bytes, size_t block_bytes, int reps, double* gbps);

// __gnu_parallel::multiway_merge of k sorted runs of `run_len` keys each,
// laid out back to back in `in`, into `out` (reference
// src/gowanlock_sort.cuh:91-92), with `threads` threads.
int mgh_multiway_merge(const char* dtype, const void* in, size_t k, size_t run_len, void* out, int threads);

// The segmented sort's tuning table (written by mgsort.tuning; no reference
// counterpart): loads `path` and writes the entry of the nearest size bin for
// (key_type, val_type) to *out.  Returns 0, 1 when the table has no entry for
// the pair (*out untouched), or -4 for a missing or malformed file.
int mgh_seg_tuning_lookup(const char* path, int key_type, int val_type, size_t total_keys, mgs_tuning* out);

// hipEvent_t helpers (for mgs_sort_timed).
void* mgh_event_create(void);
void mgh_event_destroy(void* e);
float mgh_event_elapsed_ms(void* a, void* b);

#ifdef __cplusplus
}
#endif
