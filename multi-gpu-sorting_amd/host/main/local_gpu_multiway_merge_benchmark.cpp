// local_gpu_multiway_merge_benchmark: the device counterpart of
// local_cpu_multiway_merge_benchmark (src/local_cpu_multiway_merge_benchmark.cu
// :14-86, no device program in the reference): one k-way merge
// (mgs_multiway_merge) of sorted sublists of the given lengths, each the
// generator's "sorted" distribution, already in device memory.  CSV line:
//   "sorted_sublist_lengths",gpu_id,merge_duration
// The timed phase is the merge call and a wait for its stream; one untimed
// merge runs first.  The output must be sorted and hold the keys of every
// sublist, or the program prints the reference's error line and exits 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "argument_limits.h"
#include "bench_common.h"
#include "data_generator.h"
#include "device_buffer.h"
#include "hip_check.h"
#include "host_vector.h"
#include "key_traits.h"
#include "time_durations.h"

using namespace mgsort;

int main(int argc, char* argv[]) {
  const char usage[] =
      "./local_gpu_multiway_merge_benchmark <sorted_sublist_lengths> <gpu_id>\n"
      "  (at most 256 sublists)\n";
  const std::vector<std::string> a(argv + 1, argv + argc);
  if (a.size() != 2) UsageExit(usage);
  std::vector<size_t> lens;
  size_t gpu = 0;
  if (!ParseCountList(a[0], &lens) || lens.size() > MGS_MULTIWAY_MAX_RUNS || !ParseCount(a[1], &gpu))
    UsageExit(usage);
  try {
    DataGenerator gen(ArgumentLimits::GetDefaultDistributionSeed());
    size_t total = 0;
    for (size_t l : lens) total += l;
    HostVector<int32_t> host(total, false);
    std::pair<uint64_t, uint64_t> before{0, 0};
    std::vector<size_t> offs;
    for (size_t i = 0, at = 0; i < lens.size(); at += lens[i], ++i) {
      gen.ComputeDistribution<int32_t>(host.data() + at, lens[i], "sorted");
      const auto f = Fingerprint(host.data() + at, lens[i]);
      before.first += f.first;
      before.second ^= f.second;
      offs.push_back(at);
    }
    HIP_CHECK(hipSetDevice((int)gpu));
    HIP_CHECK(hipDeviceSynchronize());
    Stream s;
    DeviceBuffer in(total * 4), out(total * 4),
        temp(std::max<size_t>(mgs_multiway_merge_temp_bytes(MGS_KEY_I32, lens.size(), total), 256));
    if (total) HIP_CHECK(hipMemcpy(in.get(), host.data(), total * 4, hipMemcpyHostToDevice));
    std::vector<const void*> runs;
    for (size_t o : offs) runs.push_back(in.get<int32_t>() + o);
    auto merge_once = [&]() {
      MGS_CHECK(mgs_multiway_merge(MGS_KEY_I32, runs.data(), lens.data(), lens.size(), out.get(), temp.get(),
                                   temp.bytes(), s));
    };
    merge_once();  // untimed: loads the code objects
    s.Sync();
    HIP_CHECK(hipMemset(temp.get(), 0, sizeof(uint32_t)));  // the status word
    HIP_CHECK(hipDeviceSynchronize());
    TimeDurations::Get()->Tic("merge_phase");
    merge_once();
    s.Sync();
    TimeDurations::Get()->Toc("merge_phase");
    MGS_CHECK(mgs_temp_status(temp.get(), temp.bytes(), s));
    if (total) HIP_CHECK(hipMemcpy(host.data(), out.get(), total * 4, hipMemcpyDeviceToHost));
    std::cout << "\"";
    for (size_t i = 0; i < lens.size(); ++i) std::cout << lens[i] << (i + 1 < lens.size() ? "," : "\",");
    std::cout << gpu << "," << std::fixed << std::setprecision(9) << TimeDurations::Get()->Seconds("merge_phase")
              << "\n";
    CheckOrExit(IsSortedEncoded(host.data(), total) && Fingerprint(host.data(), total) == before);
  } catch (const std::exception& e) {
    std::cout << "Error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
