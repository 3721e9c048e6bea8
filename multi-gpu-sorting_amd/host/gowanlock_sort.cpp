// HET sort on one device (see gowanlock_sort.h).
#include "gowanlock_sort.h"

#include <omp.h>

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <memory>
#include <mutex>
#include <parallel/algorithm>
#include <thread>
#include <utility>

#include "device_buffer.h"
#include "hip_check.h"
#include "key_traits.h"
#include "time_durations.h"

namespace mgsort {

GowanlockPlan PlanGowanlock(size_t n, size_t chunk_size, size_t merge_group_size, size_t g) {
  GowanlockPlan p;
  if (n == 0) return p;
  if (chunk_size == 0) chunk_size = n / g + ((n % g != 0) ? 1 : 0);
  if (merge_group_size == 0) merge_group_size = g;
  const size_t per_group = chunk_size * g;
  const size_t groups = n / per_group + (n % per_group != 0 ? 1 : 0);
  // one chunk per GPU and group; the last group of one GPU has no fillers
  const size_t total = groups * g;
  if (merge_group_size > total) merge_group_size = 1;
  size_t to_merge = total / merge_group_size - ((total % merge_group_size == 0 && merge_group_size > 1) ? 1 : 0);
  if (merge_group_size == 1) to_merge = 0;  // D1: every run goes to the final merge
  p.chunk_size = chunk_size;
  p.merge_group_size = merge_group_size;
  p.num_chunks = total;
  p.num_merge_groups_to_merge = to_merge;
  p.num_chunks_to_merge = to_merge * merge_group_size;
  p.final_runs = to_merge + (total - p.num_chunks_to_merge);
  return p;
}

void SetupGowanlockSort(const std::vector<int>& gpus, int num_threads) {
  if (gpus.size() != 1) throw Error("the HET sort of this build drives one GPU");
  omp_set_num_threads(num_threads);
  HIP_CHECK(hipSetDevice(gpus.front()));
  HIP_CHECK(hipDeviceSynchronize());
}

namespace {

// Sorted-run counter shared with the merge thread (SyncState of the
// reference, src/gowanlock_sort.cuh:20-27).
struct Progress {
  std::mutex mu;
  std::condition_variable cv;
  size_t sorted = 0;
};

template <typename T>
using Run = std::pair<T*, T*>;

template <typename T>
void MultiwayMerge(std::vector<Run<T>>& runs, T* out, size_t len) {
  __gnu_parallel::multiway_merge(runs.begin(), runs.end(), out, (std::ptrdiff_t)len, EncodedLess<T>());
}

// Pairwise merge-path rounds over sorted runs resident in `a` (scratch `b`);
// returns the buffer holding the merged result.
template <typename T>
T* MergeRunsOnDevice(std::vector<std::pair<size_t, size_t>> runs, T* a, T* b, hipStream_t s) {
  const mgs_key kt = KeyOf<T>::value;
  size_t longest = 0;
  for (auto& r : runs) longest = std::max(longest, r.second);
  DeviceBuffer temp;
  while (runs.size() > 1) {
    std::vector<std::pair<size_t, size_t>> next;
    for (size_t i = 0; i < runs.size(); i += 2) {
      if (i + 1 == runs.size()) {
        HIP_CHECK(hipMemcpyAsync(b + runs[i].first, a + runs[i].first, runs[i].second * sizeof(T),
                                 hipMemcpyDeviceToDevice, s));
        next.push_back(runs[i]);
        continue;
      }
      const auto& x = runs[i];
      const auto& y = runs[i + 1];
      const size_t need = mgs_merge_temp_bytes(kt, MGS_VAL_NONE, x.second, y.second);
      if (temp.bytes() < need) {
        HIP_CHECK(hipStreamSynchronize(s));
        temp.Allocate(need);
      }
      MGS_CHECK(mgs_merge(kt, MGS_VAL_NONE, a + x.first, nullptr, x.second, a + y.first, nullptr, y.second,
                          b + x.first, nullptr, temp.get(), temp.bytes(), s));
      next.push_back({x.first, x.second + y.second});
    }
    runs.swap(next);
    std::swap(a, b);
  }
  HIP_CHECK(hipStreamSynchronize(s));
  return a;
}

}  // namespace

template <typename T>
void GowanlockSort(HostVector<T>* elements, size_t chunk_size, size_t merge_group_size, size_t num_buffers,
                   const std::vector<int>& gpus, const std::string& merge_backend) {
  if (elements->empty()) return;
  if (num_buffers < 2 || num_buffers > 3) throw Error("num_buffers must be 2 or 3");
  const size_t n = elements->size();
  const GowanlockPlan plan = PlanGowanlock(n, chunk_size, merge_group_size, gpus.size());
  const size_t C = plan.num_chunks, cs = plan.chunk_size, mgs = plan.merge_group_size;
  const mgs_key kt = KeyOf<T>::value;
  auto chunk_len = [&](size_t c) { return std::min(cs, n - c * cs); };

  bool gpu_merge = merge_backend == "gpu" && C > 1;
  if (gpu_merge) {
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    const size_t need = 2 * n * sizeof(T) + mgs_sort_temp_bytes(kt, MGS_VAL_NONE, cs) + (size_t(1) << 28);
    gpu_merge = need < free_b;
  }

  TimeDurations::Get()->Tic("memory_allocation");
  HostVector<T> secondary;
  if (!gpu_merge && plan.final_runs > 1) secondary.Resize(n, false);
  const size_t nb = gpu_merge ? 1 : num_buffers;
  std::vector<DeviceBuffer> bufs(nb);
  DeviceBuffer resident, resident_alt;
  if (gpu_merge) {
    resident.Allocate(n * sizeof(T));
    resident_alt.Allocate(n * sizeof(T));
  } else {
    for (auto& b : bufs) b.Allocate(cs * sizeof(T));
  }
  DeviceBuffer sort_temp(mgs_sort_temp_bytes(kt, MGS_VAL_NONE, cs));
  Stream h2d, comp, d2h;
  std::vector<std::unique_ptr<Event>> h2d_done, sort_done, d2h_done;
  for (size_t c = 0; c < C; ++c) {
    h2d_done.emplace_back(new Event());
    sort_done.emplace_back(new Event());
    d2h_done.emplace_back(new Event());
  }
  TimeDurations::Get()->Toc("memory_allocation");

  TimeDurations::Get()->Tic("sort_phase");
  T* primary = elements->data();
  Progress prog;
  std::thread merger;
  if (!gpu_merge && plan.num_merge_groups_to_merge > 0) {
    // MergeChunks: merge each group of mgs sorted runs into `secondary` as
    // soon as its last run has landed
    merger = std::thread([&] {
      for (size_t gidx = 0; gidx < plan.num_merge_groups_to_merge; ++gidx) {
        {
          std::unique_lock<std::mutex> lk(prog.mu);
          prog.cv.wait(lk, [&] { return prog.sorted >= (gidx + 1) * mgs; });
        }
        std::vector<Run<T>> runs;
        size_t len = 0;
        const size_t first = gidx * mgs;
        for (size_t c = first; c < first + mgs; ++c) {
          runs.push_back({primary + c * cs, primary + c * cs + chunk_len(c)});
          len += chunk_len(c);
        }
        MultiwayMerge(runs, secondary.data() + first * cs, len);
      }
    });
  }

  // Chunk pipeline: H2D(c) waits until D2H(c - nb) freed its buffer.
  for (size_t c = 0; c < C; ++c) {
    const size_t len = chunk_len(c), off = c * cs;
    T* dbuf = gpu_merge ? resident.get<T>() + off : bufs[c % nb].get<T>();
    if (!gpu_merge && c >= nb) HIP_CHECK(hipStreamWaitEvent(h2d, *d2h_done[c - nb], 0));
    HIP_CHECK(hipMemcpyAsync(dbuf, primary + off, len * sizeof(T), hipMemcpyHostToDevice, h2d));
    h2d_done[c]->Record(h2d);
    HIP_CHECK(hipStreamWaitEvent(comp, *h2d_done[c], 0));
    MGS_CHECK(mgs_sort(kt, MGS_VAL_NONE, dbuf, dbuf, nullptr, nullptr, len, sort_temp.get(), sort_temp.bytes(), comp));
    sort_done[c]->Record(comp);
    if (!gpu_merge) {
      T* dst = (c < plan.num_chunks_to_merge || plan.final_runs == 1) ? primary + off : secondary.data() + off;
      HIP_CHECK(hipStreamWaitEvent(d2h, *sort_done[c], 0));
      HIP_CHECK(hipMemcpyAsync(dst, dbuf, len * sizeof(T), hipMemcpyDeviceToHost, d2h));
      d2h_done[c]->Record(d2h);
    }
  }
  if (gpu_merge) {
    comp.Sync();
    TimeDurations::Get()->Tic("merge_phase");
    std::vector<std::pair<size_t, size_t>> runs;
    for (size_t c = 0; c < C; ++c) runs.push_back({c * cs, chunk_len(c)});
    T* res = MergeRunsOnDevice(runs, resident.get<T>(), resident_alt.get<T>(), comp.get());
    HIP_CHECK(hipMemcpyAsync(primary, res, n * sizeof(T), hipMemcpyDeviceToHost, comp));
    comp.Sync();
    TimeDurations::Get()->Toc("merge_phase");
  } else {
    for (size_t c = 0; c < C; ++c) {
      d2h_done[c]->Sync();
      {
        std::lock_guard<std::mutex> lk(prog.mu);
        prog.sorted = c + 1;
      }
      prog.cv.notify_one();
    }
    if (merger.joinable()) merger.join();
    TimeDurations::Get()->Tic("merge_phase");
    if (plan.final_runs > 1) {
      // MergeMergeGroups: the pre-merged groups and the remaining runs, all
      // in `secondary`, into the caller's array
      std::vector<Run<T>> runs;
      for (size_t gidx = 0; gidx < plan.num_merge_groups_to_merge; ++gidx) {
        const size_t first = gidx * mgs;
        size_t len = 0;
        for (size_t c = first; c < first + mgs; ++c) len += chunk_len(c);
        runs.push_back({secondary.data() + first * cs, secondary.data() + first * cs + len});
      }
      for (size_t c = plan.num_chunks_to_merge; c < C; ++c)
        runs.push_back({secondary.data() + c * cs, secondary.data() + c * cs + chunk_len(c)});
      MultiwayMerge(runs, primary, n);
    }
    TimeDurations::Get()->Toc("merge_phase");
  }
  TimeDurations::Get()->Toc("sort_phase");

  TimeDurations::Get()->Tic("memory_deallocation");
  bufs.clear();
  resident.Free();
  resident_alt.Free();
  sort_temp.Free();
  secondary.Free();
  TimeDurations::Get()->Toc("memory_deallocation");
}

template void GowanlockSort<int32_t>(HostVector<int32_t>*, size_t, size_t, size_t, const std::vector<int>&,
                                     const std::string&);
template void GowanlockSort<int64_t>(HostVector<int64_t>*, size_t, size_t, size_t, const std::vector<int>&,
                                     const std::string&);
template void GowanlockSort<uint32_t>(HostVector<uint32_t>*, size_t, size_t, size_t, const std::vector<int>&,
                                      const std::string&);
template void GowanlockSort<uint64_t>(HostVector<uint64_t>*, size_t, size_t, size_t, const std::vector<int>&,
                                      const std::string&);
template void GowanlockSort<float>(HostVector<float>*, size_t, size_t, size_t, const std::vector<int>&,
                                   const std::string&);
template void GowanlockSort<double>(HostVector<double>*, size_t, size_t, size_t, const std::vector<int>&,
                                    const std::string&);

}  // namespace mgsort
