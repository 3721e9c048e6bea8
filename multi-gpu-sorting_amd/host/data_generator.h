// Seeded input generator of the sort benchmarks.
//
// Reproduces the reference's DataGenerator (src/data_generator.cuh:12-211)
// bit for bit on x86-64 libstdc++: 128 logical generator threads, thread t
// owns the t-th block of libgomp's static schedule and draws from a
// std::mt19937 seeded by SeedRandomGenerator(seed + t) (src/data_generator.cuh
// :193-205).  The 128 logical streams are independent of how many cores run
// them, so the data does not change with OMP_NUM_THREADS or the box.
//
// Reference distributions: uniform, normal, zero, staggered, sorted,
// reverse-sorted, nearly-sorted, bucket-sorted (:17-36).  Deviations, all
// where the reference is undefined or racy:
//   - nearly-sorted reads every gap from the sorted array (the reference's
//     threads race on their last element, :140-159);
//   - staggered with fewer than 10 elements is generated as uniform (the
//     reference divides by zero, :82-92);
//   - out-of-range double -> integer conversions follow x86 cvttsd2si for the
//     reference's int/long (INT_MIN / LONG_MIN) and saturate for the added
//     unsigned types.
//
// Added distributions (configs of the MI355X benchmarks; counter-based
// splitmix64 streams, so any slice can be generated independently):
//   normal-clamped   normal(mean 2^(w-1), sd 2^(w-4)) clamped to the type
//   zipf             zipf(s = 1.0) over 2^20 distinct values mapped to keys
//                    by a seeded hash
//   zipf-1.2         the same with s = 1.2
//   nearly-sorted-swaps   sorted uniform keys, then n/100 seeded
//                    transpositions applied in order
//   float-specials   uniform keys with 1% of positions replaced by +-0,
//                    +-inf, NaNs with payloads and denormals of both signs
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace mgsort {

class DataGenerator {
 public:
  static constexpr int kNumThreads = 128;

  explicit DataGenerator(uint32_t distribution_seed, int num_threads = kNumThreads)
      : distribution_seed_(distribution_seed), num_threads_(num_threads) {}

  // Fills begin[0, n).  Returns false for an unknown distribution name.
  template <typename T>
  bool ComputeDistribution(T* begin, size_t n, const std::string& distribution_type) const;

  static const std::vector<std::string>& ReferenceDistributions();
  static const std::vector<std::string>& AddedDistributions();
  static bool IsValidDistribution(const std::string& name);

  // [b, e) of logical thread t under libgomp's static schedule of n
  // iterations over `threads` threads.
  static void StaticRange(size_t n, int threads, int t, size_t* b, size_t* e);

 private:
  const size_t distribution_seed_;
  const int num_threads_;
};

}  // namespace mgsort
