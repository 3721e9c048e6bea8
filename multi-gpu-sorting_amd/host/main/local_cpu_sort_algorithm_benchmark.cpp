// local_cpu_sort_algorithm_benchmark: one host sort of the reference's
// default input, timed (src/local_cpu_sort_algorithm_benchmark.cu:13-61).
// Same arguments and CSV line:
//   num_elements,"sort_algorithm",num_threads,local_sort_duration
// Algorithms: gnu_parallel::sort, paradis::sort (this library's PARADIS),
// std::sort (one thread).  boost::sort::block_indirect_sort is named by the
// reference but Boost is not part of this build: asking for it is an error.
// The output must be sorted and a permutation of the input, or the program
// prints the reference's error line and exits 1.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <parallel/algorithm>
#include <string>
#include <vector>

#include "argument_limits.h"
#include "bench_common.h"
#include "data_generator.h"
#include "paradis.h"
#include "time_durations.h"

using namespace mgsort;

int main(int argc, char* argv[]) {
  const char usage[] =
      "./local_cpu_sort_algorithm_benchmark <num_elements> <sort_algorithm> (num_threads)\n"
      "  sort_algorithm: gnu_parallel::sort, paradis::sort, std::sort\n";
  const std::vector<std::string> a(argv + 1, argv + argc);
  if (a.size() < 2 || a.size() > 3) UsageExit(usage);
  size_t n = 0, threads = (size_t)ArgumentLimits::GetDefaultNumThreads();
  if (!ParseCount(a[0], &n) || (a.size() == 3 && (!ParseCount(a[2], &threads) || threads == 0))) UsageExit(usage);
  const std::string& alg = a[1];
  if (alg == "boost::sort::block_indirect_sort") {
    std::cout << "Error: boost::sort::block_indirect_sort is not available (Boost is not part of this build).\n";
    return 1;
  }
  if (alg != "gnu_parallel::sort" && alg != "paradis::sort" && alg != "std::sort") UsageExit(usage);
  omp_set_num_threads((int)threads);

  DataGenerator gen(ArgumentLimits::GetDefaultDistributionSeed());
  std::vector<int> e(n);
  gen.ComputeDistribution<int>(e.data(), n, ArgumentLimits::GetDefaultDistributionType());
  const auto before = Fingerprint(e.data(), n);

  TimeDurations::Get()->Tic("sort_phase");
  if (alg == "gnu_parallel::sort") __gnu_parallel::sort(e.begin(), e.end());
  else if (alg == "paradis::sort") ParadisSort(e.data(), e.data() + n, (int)threads);
  else std::sort(e.begin(), e.end());
  TimeDurations::Get()->Toc("sort_phase");

  std::cout << n << ",\"" << alg << "\"," << threads << "," << std::fixed << std::setprecision(9)
            << TimeDurations::Get()->Seconds("sort_phase") << "\n";
  CheckOrExit(IsSortedEncoded(e.data(), n) && Fingerprint(e.data(), n) == before);
  return 0;
}
