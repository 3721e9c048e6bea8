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

#include "chunk_merger.h"
#include "device_buffer.h"
#include "hip_check.h"
#include "inplace_transfer.h"
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

// Stretches of the device merge's output copied back one after another.
constexpr size_t kMergePortions = 16;

template <typename T>
using Run = std::pair<T*, T*>;

}  // namespace

template <typename T>
void GowanlockSort(HostVector<T>* elements, size_t chunk_size, size_t merge_group_size, size_t num_buffers,
                   const std::vector<int>& gpus, const std::string& merge_backend, size_t block_bytes) {
  if (elements->empty()) return;
  if (num_buffers < 2 || num_buffers > 3) throw Error("num_buffers must be 2 or 3");
  if (!elements->pinned())
    throw Error("the HET sort needs the input in pinned host memory (the pinned allocation failed)");
  const size_t n = elements->size();
  const GowanlockPlan plan = PlanGowanlock(n, chunk_size, merge_group_size, gpus.size());
  const size_t C = plan.num_chunks, cs = plan.chunk_size, mgs = plan.merge_group_size;
  const mgs_key kt = KeyOf<T>::value;
  auto chunk_len = [&](size_t c) { return std::min(cs, n - c * cs); };

  bool gpu_merge = merge_backend == "gpu" && C > 1 && C <= MGS_MULTIWAY_MAX_RUNS;
  const size_t merge_temp_bytes = gpu_merge ? mgs_multiway_merge_temp_bytes(kt, C, n) : 0;
  if (gpu_merge) {
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    const size_t need = 2 * n * sizeof(T) + std::max(mgs_sort_temp_bytes(kt, MGS_VAL_NONE, cs), merge_temp_bytes) +
                        (size_t(1) << 28);
    gpu_merge = need < free_b;
  }
  const size_t temp_bytes = mgs_sort_temp_bytes(kt, MGS_VAL_NONE, cs);

  TimeDurations::Get()->Tic("memory_allocation");
  HostVector<T> secondary;
  if (!gpu_merge && plan.final_runs > 1) {
    secondary.Resize(n, false);
    if (!secondary.pinned()) throw Error("the HET sort's secondary host buffer could not be pinned");
  }
  // Device memory plans of the reference (src/gowanlock_sort.cuh:127-393):
  //   2 buffers  two chunk buffers; the sort of the chunk in one uses the
  //              other as its scratch (2n)
  //   3 buffers  two chunk buffers plus the sort's own scratch (3n)
  //   gpu merge  every run resident, plus the merge's second copy
  std::vector<DeviceBuffer> bufs(gpu_merge ? 0 : 2);
  DeviceBuffer resident, resident_alt, sort_temp;
  if (gpu_merge) {
    // the sorts' scratch serves the merge afterwards
    resident.Allocate(n * sizeof(T));
    resident_alt.Allocate(n * sizeof(T));
    sort_temp.Allocate(std::max(temp_bytes, merge_temp_bytes));
  } else if (num_buffers == 2) {
    for (auto& b : bufs) b.Allocate(std::max(cs * sizeof(T), temp_bytes));
  } else {
    for (auto& b : bufs) b.Allocate(cs * sizeof(T));
    sort_temp.Allocate(temp_bytes);
  }
  Stream h2d, comp, d2h;
  std::vector<std::unique_ptr<Event>> h2d_done, sort_done, d2h_done;
  for (size_t c = 0; c < C; ++c) {
    h2d_done.emplace_back(new Event());
    sort_done.emplace_back(new Event());
    d2h_done.emplace_back(new Event());
  }
  // in-place transfer blocks (3 buffers): the reference moves a chunk in 100
  // blocks (InplaceMemcpy, src/gowanlock_sort.cuh:323)
  const size_t chunk_bytes = cs * sizeof(T);
  size_t block = block_bytes ? block_bytes : std::max<size_t>(1, chunk_bytes / 100);
  block = std::max<size_t>(sizeof(T), block / sizeof(T) * sizeof(T));
  const size_t max_blocks = (chunk_bytes + block - 1) / block;
  std::vector<std::unique_ptr<Event>> block_out;
  if (!gpu_merge && num_buffers == 3)
    for (size_t k = 0; k < max_blocks; ++k) block_out.emplace_back(new Event());
  TimeDurations::Get()->Toc("memory_allocation");

  TimeDurations::Get()->Tic("sort_phase");
  T* primary = elements->data();
  auto run_dst = [&](size_t c) {
    return (c < plan.num_chunks_to_merge || plan.final_runs == 1) ? primary + c * cs : secondary.data() + c * cs;
  };
  // MergeChunks: merges each group of mgs sorted runs into `secondary` as
  // soon as its last run has landed (joined, on an exception, by its
  // destructor)
  ChunkMerger<T> merger(primary, secondary.data(), cs, n, gpu_merge ? 0 : plan.num_merge_groups_to_merge, mgs);

  if (gpu_merge) {
    for (size_t c = 0; c < C; ++c) {
      const size_t len = chunk_len(c), off = c * cs;
      T* dbuf = resident.get<T>() + off;
      HIP_CHECK(hipMemcpyAsync(dbuf, primary + off, len * sizeof(T), hipMemcpyHostToDevice, h2d));
      h2d_done[c]->Record(h2d);
      HIP_CHECK(hipStreamWaitEvent(comp, *h2d_done[c], 0));
      MGS_CHECK(mgs_sort(kt, MGS_VAL_NONE, dbuf, dbuf, nullptr, nullptr, len, sort_temp.get(), sort_temp.bytes(), comp));
    }
  } else if (num_buffers == 2) {
    // GowanlockCore2n (reference :127-230): H2D(c+1) overlaps D2H(c); the
    // sort of chunk c waits until chunk c-1 has left the other buffer, which
    // it uses as scratch
    for (size_t c = 0; c < C; ++c) {
      const size_t len = chunk_len(c), off = c * cs;
      DeviceBuffer& X = bufs[c % 2];
      DeviceBuffer& Y = bufs[(c + 1) % 2];
      if (c >= 1) HIP_CHECK(hipStreamWaitEvent(h2d, *sort_done[c - 1], 0));  // X was its scratch
      if (c >= 2) HIP_CHECK(hipStreamWaitEvent(h2d, *d2h_done[c - 2], 0));  // X held chunk c-2
      HIP_CHECK(hipMemcpyAsync(X.get(), primary + off, len * sizeof(T), hipMemcpyHostToDevice, h2d));
      h2d_done[c]->Record(h2d);
      HIP_CHECK(hipStreamWaitEvent(comp, *h2d_done[c], 0));
      if (c >= 1) HIP_CHECK(hipStreamWaitEvent(comp, *d2h_done[c - 1], 0));
      MGS_CHECK(mgs_sort(kt, MGS_VAL_NONE, X.get(), X.get(), nullptr, nullptr, len, Y.get(), Y.bytes(), comp));
      sort_done[c]->Record(comp);
      HIP_CHECK(hipStreamWaitEvent(d2h, *sort_done[c], 0));
      HIP_CHECK(hipMemcpyAsync(run_dst(c), X.get(), len * sizeof(T), hipMemcpyDeviceToHost, d2h));
      d2h_done[c]->Record(d2h);
    }
  } else {
    // GowanlockCore3n (reference :232-393) with InplaceMemcpy
    // (src/inplace_memcpy.cuh:7-69): while chunk c sorts in one buffer, the
    // other buffer sends sorted chunk c-1 back block by block, and each block
    // it has sent is refilled with the same block of chunk c+1 -- the D2H of
    // block k+1 overlaps the H2D of block k, ordered by events, not by host
    // waits
    HIP_CHECK(hipMemcpyAsync(bufs[0].get(), primary, chunk_len(0) * sizeof(T), hipMemcpyHostToDevice, h2d));
    h2d_done[0]->Record(h2d);
    for (size_t c = 0; c < C; ++c) {
      DeviceBuffer& X = bufs[c % 2];
      DeviceBuffer& Y = bufs[(c + 1) % 2];
      HIP_CHECK(hipStreamWaitEvent(comp, *h2d_done[c], 0));
      MGS_CHECK(mgs_sort(kt, MGS_VAL_NONE, X.get(), X.get(), nullptr, nullptr, chunk_len(c), sort_temp.get(),
                         sort_temp.bytes(), comp));
      sort_done[c]->Record(comp);
      char* y = static_cast<char*>(Y.get());
      const size_t out_bytes = c >= 1 ? chunk_len(c - 1) * sizeof(T) : 0;
      const size_t in_bytes = c + 1 < C ? chunk_len(c + 1) * sizeof(T) : 0;
      const char* in_src = c + 1 < C ? reinterpret_cast<const char*>(primary + (c + 1) * cs) : nullptr;
      char* out_dst = c >= 1 ? reinterpret_cast<char*>(run_dst(c - 1)) : nullptr;
      if (c >= 1) HIP_CHECK(hipStreamWaitEvent(d2h, *sort_done[c - 1], 0));
      EnqueueInplaceTransfer(y, in_src, in_bytes, out_dst, out_bytes, block, h2d, d2h, block_out);
      if (c >= 1) d2h_done[c - 1]->Record(d2h);
      if (c + 1 < C) {
        // chunk c+1 may be longer than chunk c-1 was: its blocks past the
        // end of c-1 still must not overtake c-1's last block
        if (c >= 1) HIP_CHECK(hipStreamWaitEvent(h2d, *d2h_done[c - 1], 0));
        h2d_done[c + 1]->Record(h2d);
      }
    }
    HIP_CHECK(hipStreamWaitEvent(d2h, *sort_done[C - 1], 0));
    HIP_CHECK(hipMemcpyAsync(run_dst(C - 1), bufs[(C - 1) % 2].get(), chunk_len(C - 1) * sizeof(T),
                             hipMemcpyDeviceToHost, d2h));
    d2h_done[C - 1]->Record(d2h);
  }

  if (gpu_merge) {
    comp.Sync();
    TimeDurations::Get()->Tic("merge_phase");
    // one k-way merge of the resident runs (mgs_multiway_merge), written in
    // stretches that go back to the host while the next ones merge
    std::vector<const void*> runs(C);
    std::vector<size_t> lens(C);
    for (size_t c = 0; c < C; ++c) {
      runs[c] = resident.get<T>() + c * cs;
      lens[c] = chunk_len(c);
    }
    const size_t portions = std::min<size_t>(kMergePortions, std::max<size_t>(1, n / (size_t(1) << 20)));
    std::vector<size_t> ends(portions);
    std::vector<std::unique_ptr<Event>> done;
    std::vector<void*> evs;
    for (size_t p = 0; p < portions; ++p) {
      done.emplace_back(new Event());
      evs.push_back(done.back()->get());
    }
    T* merged = resident_alt.get<T>();
    MGS_CHECK(mgs_multiway_merge_portions(kt, runs.data(), lens.data(), C, merged, sort_temp.get(), sort_temp.bytes(),
                                          comp, portions, ends.data(), evs.data()));
    size_t at = 0;
    for (size_t p = 0; p < portions; ++p) {
      HIP_CHECK(hipStreamWaitEvent(d2h, *done[p], 0));
      if (ends[p] > at)
        HIP_CHECK(hipMemcpyAsync(primary + at, merged + at, (ends[p] - at) * sizeof(T), hipMemcpyDeviceToHost, d2h));
      at = ends[p];
    }
    d2h.Sync();
    MGS_CHECK(mgs_temp_status(sort_temp.get(), sort_temp.bytes(), comp));
    TimeDurations::Get()->Toc("merge_phase");
  } else {
    for (size_t c = 0; c < C; ++c) {
      d2h_done[c]->Sync();
      merger.Landed(c + 1);
    }
    merger.Join();
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
                                     const std::string&, size_t);
template void GowanlockSort<int64_t>(HostVector<int64_t>*, size_t, size_t, size_t, const std::vector<int>&,
                                     const std::string&, size_t);
template void GowanlockSort<uint32_t>(HostVector<uint32_t>*, size_t, size_t, size_t, const std::vector<int>&,
                                      const std::string&, size_t);
template void GowanlockSort<uint64_t>(HostVector<uint64_t>*, size_t, size_t, size_t, const std::vector<int>&,
                                      const std::string&, size_t);
template void GowanlockSort<float>(HostVector<float>*, size_t, size_t, size_t, const std::vector<int>&,
                                   const std::string&, size_t);
template void GowanlockSort<double>(HostVector<double>*, size_t, size_t, size_t, const std::vector<int>&,
                                    const std::string&, size_t);

}  // namespace mgsort
