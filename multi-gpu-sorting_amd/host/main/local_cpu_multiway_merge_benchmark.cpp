// local_cpu_multiway_merge_benchmark: __gnu_parallel::multiway_merge of
// sorted sublists of the given lengths (each the generator's "sorted"
// distribution), timed (src/local_cpu_multiway_merge_benchmark.cu:14-86).
// Same arguments and CSV line:
//   "sorted_sublist_lengths",num_threads,merge_duration
// The output must be sorted and hold the keys of every sublist, or the
// program prints the reference's error line and exits 1.
#include <omp.h>

#include <cstdint>
#include <functional>
#include <iomanip>
#include <iostream>
#include <parallel/algorithm>
#include <string>
#include <utility>
#include <vector>

#include "argument_limits.h"
#include "bench_common.h"
#include "data_generator.h"
#include "time_durations.h"

using namespace mgsort;

int main(int argc, char* argv[]) {
  const char usage[] = "./local_cpu_multiway_merge_benchmark <sorted_sublist_lengths> (num_threads)\n";
  const std::vector<std::string> a(argv + 1, argv + argc);
  if (a.empty() || a.size() > 2) UsageExit(usage);
  std::vector<size_t> lens;
  size_t threads = (size_t)ArgumentLimits::GetDefaultNumThreads();
  if (!ParseCountList(a[0], &lens) || (a.size() == 2 && (!ParseCount(a[1], &threads) || threads == 0)))
    UsageExit(usage);
  omp_set_num_threads((int)threads);

  DataGenerator gen(ArgumentLimits::GetDefaultDistributionSeed());
  std::vector<std::vector<int>> lists(lens.size());
  size_t total = 0;
  std::pair<uint64_t, uint64_t> before{0, 0};
  for (size_t i = 0; i < lens.size(); ++i) {
    lists[i].resize(lens[i]);
    gen.ComputeDistribution<int>(lists[i].data(), lens[i], "sorted");
    const auto f = Fingerprint(lists[i].data(), lens[i]);
    before.first += f.first;
    before.second ^= f.second;
    total += lens[i];
  }
  std::vector<int> merged(total);
  std::vector<std::pair<int*, int*>> runs;
  for (auto& l : lists) runs.emplace_back(l.data(), l.data() + l.size());

  TimeDurations::Get()->Tic("merge_phase");
  __gnu_parallel::multiway_merge(runs.begin(), runs.end(), merged.data(), (std::ptrdiff_t)total, std::less<int>());
  TimeDurations::Get()->Toc("merge_phase");

  std::cout << "\"";
  for (size_t i = 0; i < lens.size(); ++i) std::cout << lens[i] << (i + 1 < lens.size() ? "," : "\",");
  std::cout << threads << "," << std::fixed << std::setprecision(9) << TimeDurations::Get()->Seconds("merge_phase")
            << "\n";
  CheckOrExit(IsSortedEncoded(merged.data(), total) && Fingerprint(merged.data(), total) == before);
  return 0;
}
