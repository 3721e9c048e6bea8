// P2P sort on virtual partitions of one device (see tanasic_sort.h).
#include "tanasic_sort.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <memory>

#include "device_buffer.h"
#include "hip_check.h"
#include "key_traits.h"
#include "time_durations.h"

namespace mgsort {

namespace {

template <typename T>
class Partitions {
 public:
  Partitions(size_t count, size_t size) : count_(count), size_(size), cur_(count), alt_(count) {
    const mgs_key kt = KeyOf<T>::value;
    a_.Allocate(count * size * sizeof(T));
    b_.Allocate(count * size * sizeof(T));
    sort_temp_.Allocate(mgs_sort_temp_bytes(kt, MGS_VAL_NONE, size));
    merge_temp_.Allocate(mgs_merge_temp_bytes(kt, MGS_VAL_NONE, size, 0));
    for (size_t i = 0; i < count; ++i) {
      cur_[i] = a_.get<T>() + i * size;
      alt_[i] = b_.get<T>() + i * size;
    }
  }
  T* Cur(size_t i) { return cur_[i]; }
  T* Alt(size_t i) { return alt_[i]; }
  void Flip(size_t i) { std::swap(cur_[i], alt_[i]); }
  size_t size() const { return size_; }
  size_t count() const { return count_; }
  T* Base() { return a_.get<T>(); }  // contiguous view, valid while no flips are pending
  void* SortTemp() { return sort_temp_.get(); }
  size_t SortTempBytes() const { return sort_temp_.bytes(); }
  void* MergeTemp() { return merge_temp_.get(); }
  size_t MergeTempBytes() const { return merge_temp_.bytes(); }
  hipStream_t stream() const { return stream_.get(); }
  MappedSlot& slot() { return slot_; }

 private:
  size_t count_, size_;
  DeviceBuffer a_, b_, sort_temp_, merge_temp_;
  std::vector<T*> cur_, alt_;
  Stream stream_;
  MappedSlot slot_;
};

// FindPivot (reference src/tanasic_sort.cuh:45-87): p = number of elements
// the left half (ids[0, h)) and the right half (ids[h, g)) exchange.
template <typename T>
size_t FindPivot(Partitions<T>& P, const std::vector<size_t>& ids) {
  const size_t h = ids.size() / 2;
  std::vector<const void*> left(h), right(h);
  for (size_t i = 0; i < h; ++i) {
    left[i] = P.Cur(ids[i]);
    right[i] = P.Cur(ids[i + h]);
  }
  MGS_CHECK(mgs_select_pivot(KeyOf<T>::value, left.data(), right.data(), P.size(), h, P.slot().device(),
                             P.stream()));
  HIP_CHECK(hipStreamSynchronize(P.stream()));
  return (size_t)P.slot().value();
}

// SwapPartitions (reference :89-146), in place: the i-th partition left of
// the middle gives its top elements to the i-th partition right of it.
template <typename T>
std::array<size_t, 2> SwapPartitions(Partitions<T>& P, size_t pivot, const std::vector<size_t>& ids) {
  const size_t bs = P.size(), h = ids.size() / 2;
  size_t to_swap = pivot / bs;
  if (pivot == bs * h) {
    --to_swap;
    pivot = bs;
  } else {
    pivot %= bs;
  }
  std::array<size_t, 2> to_merge{};
  for (size_t i = 0; i <= to_swap; ++i) {
    const size_t l = ids[h - i - 1], r = ids[h + i];
    const size_t num = (i == to_swap) ? pivot : bs;
    MGS_CHECK(mgs_swap_ranges(KeyOf<T>::value, MGS_VAL_NONE, P.Cur(l) + bs - num, nullptr, P.Cur(r), nullptr, num,
                              P.stream()));
    if (i == to_swap) to_merge = {l, r};
  }
  return to_merge;
}

// MergeLocalPartitions (reference :148-174): the two boundary partitions
// now hold two sorted runs each; merge them (into the alternate buffer).
template <typename T>
void MergeLocalPartitions(Partitions<T>& P, size_t pivot, const std::array<size_t, 2>& to_merge,
                          const std::vector<size_t>& ids) {
  const size_t bs = P.size();
  pivot %= bs;
  for (size_t i = 0; i < ids.size(); ++i) {
    const size_t d = ids[i];
    if (d != to_merge[0] && d != to_merge[1]) continue;
    const size_t off = i >= ids.size() / 2 ? pivot : bs - pivot;
    if (off == 0 || off == bs) continue;  // one run is empty: already sorted
    MGS_CHECK(mgs_merge(KeyOf<T>::value, MGS_VAL_NONE, P.Cur(d), nullptr, off, P.Cur(d) + off, nullptr, bs - off,
                        P.Alt(d), nullptr, P.MergeTemp(), P.MergeTempBytes(), P.stream()));
    P.Flip(d);
  }
  HIP_CHECK(hipStreamSynchronize(P.stream()));
}

// MergePartitions (reference :176-204).
template <typename T>
size_t MergePartitions(Partitions<T>& P, const std::vector<size_t>& ids) {
  const size_t h = ids.size() / 2;
  if (ids.size() > 2) {
    MergePartitions(P, std::vector<size_t>(ids.begin(), ids.begin() + h));
    MergePartitions(P, std::vector<size_t>(ids.begin() + h, ids.end()));
  }
  const size_t pivot = FindPivot(P, ids);
  if (pivot > 0) {
    const auto to_merge = SwapPartitions(P, pivot, ids);
    MergeLocalPartitions(P, pivot, to_merge, ids);
  }
  if (ids.size() > 2) {
    MergePartitions(P, std::vector<size_t>(ids.begin(), ids.begin() + h));
    MergePartitions(P, std::vector<size_t>(ids.begin() + h, ids.end()));
  }
  return pivot;
}

}  // namespace

void SetupTanasicSort(const std::vector<int>& gpus) {
  if (gpus.size() != 1) throw Error("the P2P sort needs peer access between GPUs; this build runs on one GPU");
  HIP_CHECK(hipSetDevice(gpus.front()));
}

template <typename T>
size_t TanasicSort(HostVector<T>* elements, std::vector<int> gpus, size_t partitions) {
  if (elements->empty()) return 0;
  SetupTanasicSort(gpus);
  const size_t n = elements->size();
  size_t g = partitions ? partitions : gpus.size();
  TimeDurations::Get()->Tic("memory_allocation");
  // fillers and partition size as the reference (src/tanasic_sort.cuh:231-238)
  size_t fillers = (n % g != 0) ? g - n % g : 0;
  size_t bs = (n + fillers) / g;
  while (bs < fillers) {
    g /= 2;
    fillers = (n % g != 0) ? g - n % g : 0;
    bs = (n + fillers) / g;
  }
  auto owner = std::make_unique<Partitions<T>>(g, bs);
  Partitions<T>& P = *owner;
  TimeDurations::Get()->Toc("memory_allocation");

  TimeDurations::Get()->Tic("sort_phase");
  const mgs_key kt = KeyOf<T>::value;
  HIP_CHECK(hipMemcpyAsync(P.Base(), elements->data(), n * sizeof(T), hipMemcpyHostToDevice, P.stream()));
  if (fillers) MGS_CHECK(mgs_fill_max(kt, P.Base() + n, fillers, P.stream()));
  for (size_t i = 0; i < g; ++i)
    MGS_CHECK(mgs_sort(kt, MGS_VAL_NONE, P.Cur(i), P.Cur(i), nullptr, nullptr, bs, P.SortTemp(), P.SortTempBytes(),
                       P.stream()));
  HIP_CHECK(hipStreamSynchronize(P.stream()));
  TimeDurations::Get()->Tic("merge_phase");
  std::vector<size_t> ids(g);
  for (size_t i = 0; i < g; ++i) ids[i] = i;
  const size_t pivot = g > 1 ? MergePartitions(P, ids) : 0;
  TimeDurations::Get()->Toc("merge_phase");
  for (size_t i = 0; i < g; ++i) {
    const size_t cnt = (i == g - 1) ? bs - fillers : bs;
    HIP_CHECK(hipMemcpyAsync(elements->data() + i * bs, P.Cur(i), cnt * sizeof(T), hipMemcpyDeviceToHost,
                             P.stream()));
  }
  HIP_CHECK(hipStreamSynchronize(P.stream()));
  TimeDurations::Get()->Toc("sort_phase");

  TimeDurations::Get()->Tic("memory_deallocation");
  owner.reset();
  TimeDurations::Get()->Toc("memory_deallocation");
  return pivot;
}

template size_t TanasicSort<int32_t>(HostVector<int32_t>*, std::vector<int>, size_t);
template size_t TanasicSort<int64_t>(HostVector<int64_t>*, std::vector<int>, size_t);
template size_t TanasicSort<uint32_t>(HostVector<uint32_t>*, std::vector<int>, size_t);
template size_t TanasicSort<uint64_t>(HostVector<uint64_t>*, std::vector<int>, size_t);
template size_t TanasicSort<float>(HostVector<float>*, std::vector<int>, size_t);
template size_t TanasicSort<double>(HostVector<double>*, std::vector<int>, size_t);

}  // namespace mgsort
