// memory_benchmark: device allocation of num_bytes, timed
// (src/memory_benchmark.cu:11-66).  Same positional arguments and CSV line:
//   num_bytes,num_gpus,execution_mode,"memory_allocation_durations",
//   memory_allocation_total_duration
// The reference times thrust::device_vector::resize, which allocates and
// zero-fills; this times hipMalloc followed by a zero fill (hipMemset) and a
// device wait.  This build measures one GPU: num_gpus must be 1.
#include <hip/hip_runtime.h>

#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "bench_common.h"
#include "hip_check.h"
#include "time_durations.h"

using namespace mgsort;

int main(int argc, char* argv[]) {
  const char usage[] = "./memory_benchmark <num_bytes> <num_gpus> <execution_type=serial,parallel>\n";
  const std::vector<std::string> a(argv + 1, argv + argc);
  if (a.size() != 3) UsageExit(usage);
  size_t nb = 0, ngpus = 0;
  if (!ParseCount(a[0], &nb) || !ParseCount(a[1], &ngpus) || (a[2] != "serial" && a[2] != "parallel"))
    UsageExit(usage);
  if (ngpus != 1) {
    std::cout << "Error: this build measures one GPU (num_gpus 1).\n";
    return 1;
  }
  TimeDurations* td = TimeDurations::Get();
  try {
    HIP_CHECK(hipSetDevice(0));
    HIP_CHECK(hipDeviceSynchronize());
    void* p = nullptr;
    td->Tic("total_memory_allocation");
    td->Tic("memory_allocation_0");
    HIP_CHECK(hipMalloc(&p, nb));
    HIP_CHECK(hipMemset(p, 0, nb));
    HIP_CHECK(hipDeviceSynchronize());
    td->Toc("memory_allocation_0");
    td->Toc("total_memory_allocation");
    HIP_CHECK(hipFree(p));
  } catch (const std::exception& e) {
    std::cout << "Error: " << e.what() << "\n";
    return 1;
  }
  std::cout << nb << "," << ngpus << "," << a[2] << ",\"" << std::fixed << std::setprecision(9)
            << td->Seconds("memory_allocation_0") << "\"," << td->Seconds("total_memory_allocation") << "\n";
  return 0;
}
