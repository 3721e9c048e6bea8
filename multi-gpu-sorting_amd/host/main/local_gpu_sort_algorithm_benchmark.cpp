// local_gpu_sort_algorithm_benchmark: one device sort of the reference's
// default input, timed (src/local_gpu_sort_algorithm_benchmark.cu:20-97).
// Same positional arguments and CSV line:
//   num_elements,"sort_algorithm",gpu_id,local_sort_duration
// with two optional arguments after the reference's three (data type int or
// long, and a distribution of the generator; uniform by default).
//
// Algorithms (the reference's CUB, Thrust and moderngpu names select the
// ROCm library that takes their place):
//   mgs_sort                          this library's sort (the path it picks)
//   mgs_sort_lsd                      the same with the LSD passes forced
//   hipcub::DeviceRadixSort::SortKeys AMD's radix sort (hipCUB over rocPRIM);
//                                     also cub::DeviceRadixSort::SortKeys
//   thrust::sort                      rocThrust
//   hipcub::DeviceMergeSort::SortKeys a library merge sort; also
//                                     mgpu::mergesort
//
// The timed phase is the sort call and a wait for its stream, with the data
// and the scratch already on the device (the reference draws its scratch from
// a pool allocated beforehand, :38-39; thrust::sort gets a pool of 3 n keys
// + 256 MB here).  One untimed sort of the same input
// runs first, so that the time excludes loading the code objects, which
// the reference's single cold call includes.  The output must be sorted and
// a permutation of the input, or the program prints the reference's error
// line and exits 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>

#include "argument_limits.h"
#include "bench_common.h"
#include "data_generator.h"
#include "device_buffer.h"
#include "hip_check.h"
#include "host_vector.h"
#include "key_traits.h"
#include "time_durations.h"
#include "vendor_sorts.h"

using namespace mgsort;

namespace {

const char kUsage[] =
    "./local_gpu_sort_algorithm_benchmark <num_elements> <sort_algorithm> <gpu_id> [data_type] "
    "[distribution_type]\n"
    "  sort_algorithm: mgs_sort, mgs_sort_lsd, hipcub::DeviceRadixSort::SortKeys (= cub::DeviceRadixSort::SortKeys),\n"
    "                  thrust::sort, hipcub::DeviceMergeSort::SortKeys (= mgpu::mergesort)\n"
    "  data_type: int (default), long\n";

enum class Leg { kMgs, kMgsLsd, kRadix, kThrust, kMergeSort };

bool ParseLeg(const std::string& a, Leg* leg) {
  if (a == "mgs_sort") *leg = Leg::kMgs;
  else if (a == "mgs_sort_lsd") *leg = Leg::kMgsLsd;
  else if (a == "hipcub::DeviceRadixSort::SortKeys" || a == "cub::DeviceRadixSort::SortKeys") *leg = Leg::kRadix;
  else if (a == "thrust::sort") *leg = Leg::kThrust;
  else if (a == "hipcub::DeviceMergeSort::SortKeys" || a == "mgpu::mergesort") *leg = Leg::kMergeSort;
  else return false;
  return true;
}

template <typename T>
int Run(size_t n, const std::string& name, Leg leg, int gpu, const std::string& dist) {
  DataGenerator gen(ArgumentLimits::GetDefaultDistributionSeed());
  HostVector<T> elements(n, false);
  if (!gen.ComputeDistribution<T>(elements.data(), n, dist)) UsageExit(kUsage);
  const auto before = Fingerprint(elements.data(), n);

  HIP_CHECK(hipSetDevice(gpu));
  HIP_CHECK(hipDeviceSynchronize());
  Stream s;
  const mgs_key kt = KeyOf<T>::value;
  DeviceBuffer keys(n * sizeof(T)), out;
  size_t temp_bytes = 0;
  if (leg == Leg::kMgs || leg == Leg::kMgsLsd) {
    if (leg == Leg::kMgsLsd) setenv("MGS_MSD", "0", 1);
    temp_bytes = mgs_sort_temp_bytes(kt, MGS_VAL_NONE, n);
  } else if (leg == Leg::kRadix) {
    out.Allocate(n * sizeof(T));
    HIP_CHECK((hipError_t)vendor::RadixSortKeys<T>(nullptr, &temp_bytes, keys.get<T>(), out.get<T>(), n, s));
  } else if (leg == Leg::kMergeSort) {
    HIP_CHECK((hipError_t)vendor::MergeSortKeys<T>(nullptr, &temp_bytes, keys.get<T>(), n, s));
  } else {
    // rocThrust's radix sort needs more than the reference's pool of
    // n keys + 128 MB (:38-39) from 2e9 int keys on
    temp_bytes = 3 * n * sizeof(T) + 256000000;
  }
  // never empty: a null scratch pointer makes the library calls a size query
  DeviceBuffer temp(std::max<size_t>(temp_bytes, 256));

  auto sort_once = [&]() {
    if (leg == Leg::kMgs || leg == Leg::kMgsLsd) {
      MGS_CHECK(mgs_sort(kt, MGS_VAL_NONE, keys.get(), keys.get(), nullptr, nullptr, n, temp.get(), temp.bytes(), s));
    } else if (leg == Leg::kRadix) {
      size_t tb = temp.bytes();
      HIP_CHECK((hipError_t)vendor::RadixSortKeys<T>(temp.get(), &tb, keys.get<T>(), out.get<T>(), n, s));
    } else if (leg == Leg::kMergeSort) {
      size_t tb = temp.bytes();
      HIP_CHECK((hipError_t)vendor::MergeSortKeys<T>(temp.get(), &tb, keys.get<T>(), n, s));
    } else {
      HIP_CHECK((hipError_t)vendor::ThrustSort<T>(keys.get<T>(), n, temp.get(), temp.bytes(), s));
    }
  };

  if (n) HIP_CHECK(hipMemcpy(keys.get(), elements.data(), n * sizeof(T), hipMemcpyHostToDevice));
  sort_once();  // untimed: loads the code objects
  s.Sync();
  if (n) HIP_CHECK(hipMemcpy(keys.get(), elements.data(), n * sizeof(T), hipMemcpyHostToDevice));
  // the status word of mgs_sort's scratch: sorts that need no scratch leave it as it is
  HIP_CHECK(hipMemset(temp.get(), 0, sizeof(uint32_t)));
  HIP_CHECK(hipDeviceSynchronize());

  TimeDurations::Get()->Tic("local_sort_phase");
  sort_once();
  s.Sync();
  TimeDurations::Get()->Toc("local_sort_phase");
  if (leg == Leg::kMgs || leg == Leg::kMgsLsd) MGS_CHECK(mgs_temp_status(temp.get(), temp.bytes(), s));

  const void* result = leg == Leg::kRadix ? out.get() : keys.get();
  if (n) HIP_CHECK(hipMemcpy(elements.data(), result, n * sizeof(T), hipMemcpyDeviceToHost));

  std::cout << n << ",\"" << name << "\"," << gpu << "," << std::fixed << std::setprecision(9)
            << TimeDurations::Get()->Seconds("local_sort_phase") << "\n";
  CheckOrExit(IsSortedEncoded(elements.data(), n) && Fingerprint(elements.data(), n) == before);
  return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  const std::vector<std::string> a(argv + 1, argv + argc);
  if (a.size() < 3 || a.size() > 5) UsageExit(kUsage);
  size_t n = 0, gpu = 0;
  Leg leg;
  if (!ParseCount(a[0], &n) || !ParseLeg(a[1], &leg) || !ParseCount(a[2], &gpu)) UsageExit(kUsage);
  const std::string type = a.size() > 3 ? a[3] : "int";
  const std::string dist = a.size() > 4 ? a[4] : ArgumentLimits::GetDefaultDistributionType();
  try {
    if (type == "int") return Run<int32_t>(n, a[1], leg, (int)gpu, dist);
    if (type == "long") return Run<int64_t>(n, a[1], leg, (int)gpu, dist);
  } catch (const std::exception& e) {
    std::cout << "Error: " << e.what() << "\n";
    return 1;
  }
  UsageExit(kUsage);
}
