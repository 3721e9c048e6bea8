// tsan_check: the threaded host code of the library built with
// -fsanitize=thread (and the OpenMP runtime's ThreadSanitizer support), run
// on inputs small enough for the sanitizer:
//   paradis    PARADIS (OpenMP: parallel histograms, speculative permutation
//              and repair rounds, parallel recursion) on u32 / u64 / f32 keys
//              of several distributions and thread counts
//   generator  the DataGenerator's OpenMP loops (uniform, normal, zipf,
//              nearly sorted)
//   merger     the HET sort's merge thread (ChunkMerger): producer threads
//              land sorted chunks in the primary buffer in order while the
//              merge thread merges finished groups, then the final
//              multiway merge; and the exception path (the destructor wakes
//              and joins the thread)
// Every result is compared with std::sort.  Exits 0 when all checks pass
// (the sanitizer itself exits with its own code on a report).
#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "chunk_merger.h"
#include "data_generator.h"
#include "key_traits.h"
#include "paradis.h"

using namespace mgsort;

namespace {

int failures = 0;

void Check(bool ok, const std::string& what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
  std::fflush(stdout);
  if (!ok) ++failures;
}

template <typename T>
std::vector<T> Keys(size_t n, const std::string& dist, uint64_t seed) {
  std::vector<T> v(n);
  std::mt19937_64 g(seed);
  if (dist == "uniform") {
    for (auto& x : v) {
      const uint64_t r = g();
      std::memcpy(&x, &r, sizeof(T));
    }
  } else if (dist == "few") {
    for (auto& x : v) x = (T)(g() % 5);
  } else {  // sorted runs with swaps
    for (size_t i = 0; i < n; ++i) v[i] = (T)i;
    for (size_t i = 0; i < n / 100; ++i) std::swap(v[g() % n], v[g() % n]);
  }
  if constexpr (std::is_floating_point_v<T>)
    for (auto& x : v)
      if (x != x) x = T(1);  // NaN-free (PARADIS orders by the encoded image)
  return v;
}

template <typename T>
void CheckParadis(size_t n, const std::string& dist, int threads) {
  std::vector<T> a = Keys<T>(n, dist, n + threads);
  std::vector<T> b = a;
  ParadisSort(a.data(), a.data() + n, threads);
  std::sort(b.begin(), b.end(), EncodedLess<T>());
  Check(std::memcmp(a.data(), b.data(), n * sizeof(T)) == 0,
        "paradis " + std::to_string(sizeof(T) * 8) + "-bit " + dist + " n=" + std::to_string(n) +
            " threads=" + std::to_string(threads));
}

template <typename T>
void CheckGenerator(const std::string& dist, size_t n) {
  DataGenerator gen(2147483647);
  std::vector<T> a(n), b(n);
  omp_set_num_threads(8);
  const bool ok1 = gen.ComputeDistribution<T>(a.data(), n, dist);
  omp_set_num_threads(1);
  const bool ok2 = gen.ComputeDistribution<T>(b.data(), n, dist);
  // the generator's output does not depend on the thread count
  Check(ok1 && ok2 && std::memcmp(a.data(), b.data(), n * sizeof(T)) == 0, "generator " + dist);
}

// Producers land chunks 0, 1, ... in order (each sorts its chunk, then
// publishes the count), as the device pipeline's copy-back events do.
void CheckMerger(size_t n, size_t cs, size_t group) {
  // the plan of PlanGowanlock: merge groups of `group` chunks except the
  // last group, none for groups of one chunk
  const size_t chunks = (n + cs - 1) / cs;
  const size_t groups = group <= 1 ? 0 : chunks / group - (chunks % group == 0 ? 1 : 0);
  std::vector<uint32_t> primary = Keys<uint32_t>(n, "uniform", 99), secondary(n), expect = primary;
  std::sort(expect.begin(), expect.end());
  {
    ChunkMerger<uint32_t> merger(primary.data(), secondary.data(), cs, n, groups, group);
    std::mutex order_mu;
    std::condition_variable order_cv;
    size_t next = 0;
    std::vector<std::thread> producers;
    for (int p = 0; p < 3; ++p)
      producers.emplace_back([&, p] {
        for (size_t c = p; c < chunks; c += 3) {
          const size_t len = std::min(cs, n - c * cs);
          std::sort(primary.begin() + c * cs, primary.begin() + c * cs + len);
          std::unique_lock<std::mutex> lk(order_mu);
          order_cv.wait(lk, [&] { return next == c; });
          next = c + 1;
          merger.Landed(next);
          order_cv.notify_all();
        }
      });
    for (auto& t : producers) t.join();
    merger.Join();
  }
  // MergeMergeGroups: the merged groups and the remaining chunks
  std::vector<std::pair<uint32_t*, uint32_t*>> runs;
  for (size_t g = 0; g < groups; ++g) {
    size_t len = 0;
    for (size_t c = g * group; c < (g + 1) * group; ++c) len += std::min(cs, n - c * cs);
    runs.push_back({secondary.data() + g * group * cs, secondary.data() + g * group * cs + len});
  }
  for (size_t c = groups * group; c < chunks; ++c)
    runs.push_back({primary.data() + c * cs, primary.data() + c * cs + std::min(cs, n - c * cs)});
  std::vector<uint32_t> out(n);
  MultiwayMerge(runs, out.data(), n);
  Check(out == expect, "merger n=" + std::to_string(n) + " chunk=" + std::to_string(cs) + " group=" +
                           std::to_string(group) + " groups=" + std::to_string(groups));
}

void CheckMergerAbandoned() {
  // an exception before any chunk landed: the destructor wakes the thread
  std::vector<uint32_t> a = Keys<uint32_t>(1 << 16, "uniform", 5), b(1 << 16);
  for (size_t c = 0; c < 8; ++c) std::sort(a.begin() + c * 8192, a.begin() + (c + 1) * 8192);
  try {
    ChunkMerger<uint32_t> merger(a.data(), b.data(), 8192, 1 << 16, 2, 4);
    merger.Landed(3);
    throw std::runtime_error("device error");
  } catch (const std::exception&) {
  }
  Check(true, "merger exception path joins");
}

}  // namespace

int main() {
  for (int threads : {2, 4, 8}) {
    CheckParadis<uint32_t>(1 << 20, "uniform", threads);
    CheckParadis<uint32_t>(1 << 19, "few", threads);
    CheckParadis<uint64_t>(1 << 19, "uniform", threads);
    CheckParadis<uint64_t>(1 << 18, "nearly", threads);
    CheckParadis<float>(1 << 19, "uniform", threads);
  }
  for (const char* d : {"uniform", "normal", "staggered", "bucket-sorted", "nearly-sorted", "zipf", "nearly-sorted-swaps"})
    CheckGenerator<uint32_t>(d, 1 << 18);
  CheckGenerator<uint64_t>("uniform", 1 << 18);
  CheckMerger(1 << 20, 1 << 16, 4);
  CheckMerger((1 << 20) + 12345, 1 << 16, 3);
  CheckMerger(1 << 18, 1 << 14, 1);
  CheckMergerAbandoned();
  std::printf("%s (%d failed)\n", failures ? "FAILED" : "all passed", failures);
  return failures ? 1 : 0;
}
