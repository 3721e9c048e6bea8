// extern "C" surface of the host layer (libmgsort_host.so): the data
// generator, the CPU baselines, the drop-in sort drivers and event helpers,
// for callers that bind C (the Python tests and benchmarks).
#include <hip/hip_runtime.h>
#include <omp.h>

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <parallel/algorithm>
#include <string>
#include <thread>
#include <vector>

#include "data_generator.h"
#include "gowanlock_sort.h"
#include "host_api.h"
#include "host_vector.h"
#include "key_traits.h"
#include "device_buffer.h"
#include "hip_check.h"
#include "inplace_transfer.h"
#include "paradis.h"
#include "seg_tuning.h"
#include "sort_server.h"
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
                       size_t num_buffers, const char* merge_backend, int threads, size_t block_bytes) {
  return Guard([&] {
    return ByType(dtype, [&](auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      mgsort::TimeDurations::Get()->Reset();
      mgsort::SetupGowanlockSort({gpu}, threads);
      return OnHostVector<T>(keys, n, [&](mgsort::HostVector<T>& v) {
        mgsort::GowanlockSort<T>(&v, chunk_size, merge_group_size, num_buffers, {gpu},
                                 merge_backend ? merge_backend : "cpu", block_bytes);
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

int mgh_serve(mgh_serve_request* reqs, size_t count, int gpu, int batched, int max_inflight, int host_data,
              double budget_gib, double* gkeys_s) {
  if (reqs == nullptr && count > 0) return -4;
  return Guard([&] {
    using mgsort::SortServer;
    HIP_CHECK(hipSetDevice(gpu));
    mgsort::ServerConfig cfg;
    cfg.device = gpu;
    cfg.batched = batched != 0;
    cfg.max_inflight = max_inflight > 0 ? max_inflight : 1;
    cfg.host_data = host_data != 0;
    cfg.hbm_budget = (size_t)(budget_gib * (double)(size_t(1) << 30));
    if (const char* t = std::getenv("MGS_SEG_TUNING")) cfg.seg_tuning_path = t;
    SortServer server(cfg);
    auto kb = [](int kt) { return (kt == MGS_KEY_U64 || kt == MGS_KEY_I64 || kt == MGS_KEY_F64) ? size_t(8) : size_t(4); };
    auto vb = [](int vt) { return vt == MGS_VAL_U32 ? size_t(4) : vt == MGS_VAL_U64 ? size_t(8) : size_t(0); };
    std::vector<std::unique_ptr<mgsort::SortRequest>> rs(count);
    std::vector<void*> slots(count, nullptr);
    for (size_t i = 0; i < count; ++i) {
      const mgh_serve_request& q = reqs[i];
      auto r = std::make_unique<mgsort::SortRequest>();
      r->kt = (mgs_key)q.key_type;
      r->vt = (mgs_val)q.val_type;
      r->n = q.n;
      r->keys_in = q.keys_in;
      r->vals_in = q.vals_in;
      r->keys_out = q.keys_out;
      r->vals_out = q.vals_out;
      if (!host_data) {
        slots[i] = server.AllocSlot(r->kt, r->vt, q.n);
        if (!slots[i]) throw mgsort::Error("serving pool too small for the device-mode request set");
        HIP_CHECK(hipMemcpy(slots[i], q.keys_in, q.n * kb(q.key_type), hipMemcpyHostToDevice));
        if (vb(q.val_type))
          HIP_CHECK(hipMemcpy(SortServer::SlotVals(slots[i], r->kt, q.n), q.vals_in, q.n * vb(q.val_type),
                              hipMemcpyHostToDevice));
        r->slot = slots[i];
      }
      rs[i] = std::move(r);
    }
    const double t0 = SortServer::Now();
    for (size_t i = 0; i < count; ++i) {
      const double wait = t0 + reqs[i].arrival_s - SortServer::Now();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      server.Submit(rs[i].get());
    }
    server.Drain();
    double first = 1e300, last = 0;
    size_t keys = 0;
    int rc = 0;
    for (size_t i = 0; i < count; ++i) {
      mgsort::SortRequest& r = *rs[i];
      reqs[i].latency_s = r.t_done - r.t_submit;
      reqs[i].status = r.status;
      first = std::min(first, r.t_submit);
      last = std::max(last, r.t_done);
      if (r.status == MGS_OK) keys += r.n; else rc = -2;
      if (!host_data) {
        HIP_CHECK(hipMemcpy(reqs[i].keys_out, slots[i], r.n * kb(reqs[i].key_type), hipMemcpyDeviceToHost));
        if (vb(reqs[i].val_type))
          HIP_CHECK(hipMemcpy(reqs[i].vals_out, SortServer::SlotVals(slots[i], r.kt, r.n),
                              r.n * vb(reqs[i].val_type), hipMemcpyDeviceToHost));
        server.FreeSlot(slots[i]);
      }
    }
    if (gkeys_s) *gkeys_s = (count && last > first) ? keys / (last - first) / 1e9 : 0.0;
    return rc;
  });
}

int mgh_seg_tuning_lookup(const char* path, int key_type, int val_type, size_t total_keys, mgs_tuning* out) {
  if (path == nullptr || out == nullptr) return -4;
  try {
    const mgsort::SegTuningTable t = mgsort::SegTuningTable::Load(path);
    const mgs_tuning* e = t.Lookup((mgs_key)key_type, (mgs_val)val_type, total_keys);
    if (e == nullptr) return 1;
    *out = *e;
    return 0;
  } catch (const std::exception& ex) {
    g_error = ex.what();
    return -4;
  }
}

int mgh_link_probe(int gpu, int mode, size_t bytes, size_t block_bytes, int reps, double* gbps) {
  if (gbps == nullptr || bytes == 0 || mode < 0 || mode > 3) return -4;
  return Guard([&] {
    HIP_CHECK(hipSetDevice(gpu));
    mgsort::HostVector<uint8_t> h_in(bytes, false), h_out(bytes, false);
    if (!h_in.pinned() || !h_out.pinned()) throw mgsort::Error("pinned host allocation failed");
    std::memset(h_in.data(), 1, bytes);
    std::memset(h_out.data(), 2, bytes);
    mgsort::DeviceBuffer d(bytes), d2(mode == 2 ? bytes : 1);
    mgsort::Stream s1, s2;
    mgsort::Event t0(true), t1(true), ev(false);
    const size_t block = block_bytes ? std::min(block_bytes, bytes) : bytes;
    std::vector<std::unique_ptr<mgsort::Event>> out_ev;
    if (mode == 3)
      for (size_t at = 0; at < bytes; at += block) out_ev.emplace_back(new mgsort::Event());
    double best = 1e300;
    for (int r = 0; r < std::max(1, reps) + 1; ++r) {  // the first repetition warms up
      HIP_CHECK(hipDeviceSynchronize());
      t0.Record(s1);
      HIP_CHECK(hipStreamWaitEvent(s2, t0, 0));
      char* dp = static_cast<char*>(d.get());
      if (mode == 0) {
        HIP_CHECK(hipMemcpyAsync(dp, h_in.data(), bytes, hipMemcpyHostToDevice, s1));
      } else if (mode == 1) {
        HIP_CHECK(hipMemcpyAsync(h_out.data(), dp, bytes, hipMemcpyDeviceToHost, s1));
      } else if (mode == 2) {
        HIP_CHECK(hipMemcpyAsync(dp, h_in.data(), bytes, hipMemcpyHostToDevice, s1));
        HIP_CHECK(hipMemcpyAsync(h_out.data(), d2.get(), bytes, hipMemcpyDeviceToHost, s2));
      } else {
        mgsort::EnqueueInplaceTransfer(dp, reinterpret_cast<const char*>(h_in.data()), bytes,
                                       reinterpret_cast<char*>(h_out.data()), bytes, block, s1, s2, out_ev);
      }
      ev.Record(s2);
      HIP_CHECK(hipStreamWaitEvent(s1, ev, 0));
      t1.Record(s1);
      t1.Sync();
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
      if (r > 0) best = std::min(best, (double)ms);
    }
    const double moved = (mode >= 2 ? 2.0 : 1.0) * (double)bytes;
    *gbps = moved / (best / 1e3) / 1e9;
    return 0;
  });
}

int mgh_multiway_merge(const char* dtype, const void* in, size_t k, size_t run_len, void* out, int threads) {
  if ((in == nullptr || out == nullptr) && k * run_len > 0) return -4;
  return Guard([&] {
    return ByType(dtype, [&](auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      omp_set_num_threads(threads);
      // the GNU multiway merge takes mutable run iterators; it only reads them
      T* p = const_cast<T*>(static_cast<const T*>(in));
      std::vector<std::pair<T*, T*>> runs;
      for (size_t i = 0; i < k; ++i) runs.push_back({p + i * run_len, p + (i + 1) * run_len});
      __gnu_parallel::multiway_merge(runs.begin(), runs.end(), static_cast<T*>(out), (std::ptrdiff_t)(k * run_len),
                                     mgsort::EncodedLess<T>());
      return 0;
    });
  });
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
