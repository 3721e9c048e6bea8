// local_cpu_merge_algorithm_benchmark: one host merge of two sorted arrays of
// num_elements keys each (the generator's "sorted" distribution), timed
// (src/local_cpu_merge_algorithm_benchmark.cu:13-70).  Same arguments and CSV
// line:
//   num_elements,"merge_algorithm",num_threads,local_merge_duration
// Algorithms: gnu_parallel::merge, gnu_parallel::multiway_merge.  The output
// must be sorted and hold the keys of both inputs, or the program prints
// the reference's error line and exits 1.
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
  const char usage[] =
      "./local_cpu_merge_algorithm_benchmark <num_elements> <merge_algorithm> (num_threads)\n"
      "  merge_algorithm: gnu_parallel::merge, gnu_parallel::multiway_merge\n";
  const std::vector<std::string> a(argv + 1, argv + argc);
  if (a.size() < 2 || a.size() > 3) UsageExit(usage);
  size_t n = 0, threads = (size_t)ArgumentLimits::GetDefaultNumThreads();
  if (!ParseCount(a[0], &n) || (a.size() == 3 && (!ParseCount(a[2], &threads) || threads == 0))) UsageExit(usage);
  const std::string& alg = a[1];
  if (alg != "gnu_parallel::merge" && alg != "gnu_parallel::multiway_merge") UsageExit(usage);
  omp_set_num_threads((int)threads);

  DataGenerator gen(ArgumentLimits::GetDefaultDistributionSeed());
  std::vector<int> x(n), y(n), merged(2 * n);
  gen.ComputeDistribution<int>(x.data(), n, "sorted");
  gen.ComputeDistribution<int>(y.data(), n, "sorted");
  const auto fx = Fingerprint(x.data(), n), fy = Fingerprint(y.data(), n);
  const std::pair<uint64_t, uint64_t> before{fx.first + fy.first, fx.second ^ fy.second};

  TimeDurations::Get()->Tic("merge_phase");
  if (alg == "gnu_parallel::merge") {
    __gnu_parallel::merge(x.begin(), x.end(), y.begin(), y.end(), merged.begin());
  } else {
    std::vector<std::pair<int*, int*>> runs{{x.data(), x.data() + n}, {y.data(), y.data() + n}};
    __gnu_parallel::multiway_merge(runs.begin(), runs.end(), merged.data(), (std::ptrdiff_t)(2 * n), std::less<int>());
  }
  TimeDurations::Get()->Toc("merge_phase");

  std::cout << n << ",\"" << alg << "\"," << threads << "," << std::fixed << std::setprecision(9)
            << TimeDurations::Get()->Seconds("merge_phase") << "\n";
  CheckOrExit(IsSortedEncoded(merged.data(), 2 * n) && Fingerprint(merged.data(), 2 * n) == before);
  return 0;
}
