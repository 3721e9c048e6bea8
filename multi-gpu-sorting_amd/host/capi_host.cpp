// extern "C" surface of the host layer (libmgsort_host.so): the data
// generator, the CPU baselines, the drop-in sort drivers and event helpers,
// for callers that bind C (the Python tests and benchmarks).
#include <hip/hip_runtime.h>
#include <omp.h>

#include <cstring>
#include <parallel/algorithm>
#include <string>

#include "data_generator.h"
#include "gowanlock_sort.h"
#include "host_api.h"
#include "host_vector.h"
#include "key_traits.h"
#include "paradis.h"
#include "tanasic_sort.h"
#include "time_durations.h"

namespace {

thread_local std::string g_error;

template <typename Fn>
int Guard(Fn fn) {
  try {
    return fn();
  } catch (const std::exception& e) {
    g_error = e.what();
    return -3;
  }
}

// Calls fn.template operator()<T>() for the C++ type named by dtype.
template <typename Fn>
int ByType(const char* dtype, Fn fn) {
  if (dtype == nullptr) return -4;
  const std::string t(dtype);
  if (t == "int" || t == "i32") return fn((int32_t*)nullptr);
  if (t == "long" || t == "i64") return fn((int64_t*)nullptr);
  if (t == "float" || t == "f32") return fn((float*)nullptr);
  if (t == "double" || t == "f64") return fn((double*)nullptr);
  if (t == "u32") return fn((uint32_t*)nullptr);
  if (t == "u64") return fn((uint64_t*)nullptr);
  return -1;
}

// Copies `n` keys into a pinned HostVector, runs body on it, copies back.
template <typename T, typename Body>
int OnHostVector(void* keys, size_t n, Body body) {
  mgsort::HostVector<T> v(n, false);
  if (n) std::memcpy(v.data(), keys, n * sizeof(T));
  body(v);
  if (n) std::memcpy(keys, v.data(), n * sizeof(T));
  return 0;
}

}  // namespace

extern "C" {

int mgh_generate(const char* dtype, const char* dist, uint32_t seed, void* out, size_t n) {
  if (dist == nullptr || (out == nullptr && n > 0)) return -4;
  return Guard([&] {
    return ByType(dtype, [&](auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      mgsort::DataGenerator g(seed);
      return g.ComputeDistribution<T>(static_cast<T*>(out), n, dist) ? 0 : -1;
    });
  });
}

int mgh_paradis_sort(const char* dtype, void* keys, size_t n, int threads) {
  return Guard([&] {
    return ByType(dtype, [&](auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      mgsort::ParadisSort<T>(static_cast<T*>(keys), static_cast<T*>(keys) + n, threads);
      return 0;
    });
  });
}

int mgh_gnu_parallel_sort(const char* dtype, void* keys, size_t n, int threads) {
  return Guard([&] {
    return ByType(dtype, [&](auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      omp_set_num_threads(threads);
      T* p = static_cast<T*>(keys);
      __gnu_parallel::sort(p, p + n, mgsort::EncodedLess<T>());
      return 0;
    });
  });
}

int mgh_tanasic_sort(const char* dtype, void* keys, size_t n, int gpu, size_t partitions, uint64_t* pivot) {
  return Guard([&] {
    return ByType(dtype, [&](auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      mgsort::TimeDurations::Get()->Reset();
      return OnHostVector<T>(keys, n, [&](mgsort::HostVector<T>& v) {
        const size_t p = mgsort::TanasicSort<T>(&v, {gpu}, partitions);
        if (pivot) *pivot = p;
      });
    });
  });
}

int mgh_gowanlock_sort(const char* dtype, void* keys, size_t n, int gpu, size_t chunk_size, size_t merge_group_size,
                       size_t num_buffers, const char* merge_backend, int threads) {
  return Guard([&] {
    return ByType(dtype, [&](auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      mgsort::TimeDurations::Get()->Reset();
      mgsort::SetupGowanlockSort({gpu}, threads);
      return OnHostVector<T>(keys, n, [&](mgsort::HostVector<T>& v) {
        mgsort::GowanlockSort<T>(&v, chunk_size, merge_group_size, num_buffers, {gpu},
                                 merge_backend ? merge_backend : "cpu");
      });
    });
  });
}

int mgh_gowanlock_plan(size_t n, size_t chunk_size, size_t merge_group_size, size_t out[6]) {
  const mgsort::GowanlockPlan p = mgsort::PlanGowanlock(n, chunk_size, merge_group_size, 1);
  out[0] = p.chunk_size;
  out[1] = p.merge_group_size;
  out[2] = p.num_chunks;
  out[3] = p.num_merge_groups_to_merge;
  out[4] = p.num_chunks_to_merge;
  out[5] = p.final_runs;
  return 0;
}

double mgh_phase_seconds(const char* phase) { return mgsort::TimeDurations::Get()->Seconds(phase ? phase : ""); }

const char* mgh_last_error(void) { return g_error.c_str(); }

void* mgh_event_create(void) {
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void mgh_event_destroy(void* e) {
  if (e) (void)hipEventDestroy(static_cast<hipEvent_t>(e));
}

float mgh_event_elapsed_ms(void* a, void* b) {
  float ms = -1.0f;
  if (hipEventElapsedTime(&ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)) != hipSuccess) return -1.0f;
  return ms;
}

}  // extern "C"
