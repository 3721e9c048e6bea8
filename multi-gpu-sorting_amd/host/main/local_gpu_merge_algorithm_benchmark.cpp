// local_gpu_merge_algorithm_benchmark: one device merge of two sorted halves,
// timed (src/local_gpu_merge_algorithm_benchmark.cu:18-88).  Same positional
// arguments, inputs (the generator's "sorted" distribution for both halves of
// num_elements keys) and CSV line:
//   num_elements,"merge_algorithm",gpu_id,local_merge_duration
// with an optional fourth argument, the data type (int or long).
//
// Algorithms:
//   mgs_merge            this library's merge-path merge (mgs_merge); also
//                        mgpu::merge, the merge-path merge it replaces
//   mgs_multiway_merge   this library's k-way merge with k = 2
//   thrust::merge        rocThrust
//
// The timed phase is the merge call and a wait for its stream, the inputs
// and scratch already on the device; one untimed merge runs first (see
// local_gpu_sort_algorithm_benchmark).  The output must be sorted and hold
// the keys of both inputs, or the program prints the reference's error line
// and exits 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
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
    "./local_gpu_merge_algorithm_benchmark <num_elements> <merge_algorithm> <gpu_id> [data_type]\n"
    "  merge_algorithm: mgs_merge (= mgpu::merge), mgs_multiway_merge, thrust::merge\n"
    "  data_type: int (default), long\n";

enum class Leg { kMerge, kMultiway, kThrust };

bool ParseLeg(const std::string& a, Leg* leg) {
  if (a == "mgs_merge" || a == "mgpu::merge") *leg = Leg::kMerge;
  else if (a == "mgs_multiway_merge") *leg = Leg::kMultiway;
  else if (a == "thrust::merge") *leg = Leg::kThrust;
  else return false;
  return true;
}

template <typename T>
int Run(size_t n, const std::string& name, Leg leg, int gpu) {
  const size_t na = n / 2, nb = n - n / 2;
  DataGenerator gen(ArgumentLimits::GetDefaultDistributionSeed());
  HostVector<T> a(na, false), b(nb, false), merged(n, false);
  gen.ComputeDistribution<T>(a.data(), na, "sorted");
  gen.ComputeDistribution<T>(b.data(), nb, "sorted");
  auto fa = Fingerprint(a.data(), na), fb = Fingerprint(b.data(), nb);
  const std::pair<uint64_t, uint64_t> before{fa.first + fb.first, fa.second ^ fb.second};

  HIP_CHECK(hipSetDevice(gpu));
  HIP_CHECK(hipDeviceSynchronize());
  Stream s;
  const mgs_key kt = KeyOf<T>::value;
  DeviceBuffer da(na * sizeof(T)), db(nb * sizeof(T)), out(n * sizeof(T));
  const void* runs[2] = {da.get(), db.get()};
  const size_t lens[2] = {na, nb};
  size_t temp_bytes = 0;
  if (leg == Leg::kMerge) temp_bytes = mgs_merge_temp_bytes(kt, MGS_VAL_NONE, na, nb);
  else if (leg == Leg::kMultiway) temp_bytes = mgs_multiway_merge_temp_bytes(kt, 2, n);
  else temp_bytes = 128000000 + n * sizeof(T);  // the reference's pool (:51-52)
  DeviceBuffer temp(std::max<size_t>(temp_bytes, 256));
  if (na) HIP_CHECK(hipMemcpy(da.get(), a.data(), na * sizeof(T), hipMemcpyHostToDevice));
  if (nb) HIP_CHECK(hipMemcpy(db.get(), b.data(), nb * sizeof(T), hipMemcpyHostToDevice));

  auto merge_once = [&]() {
    if (leg == Leg::kMerge) {
      MGS_CHECK(mgs_merge(kt, MGS_VAL_NONE, da.get(), nullptr, na, db.get(), nullptr, nb, out.get(), nullptr,
                          temp.get(), temp.bytes(), s));
    } else if (leg == Leg::kMultiway) {
      MGS_CHECK(mgs_multiway_merge(kt, runs, lens, 2, out.get(), temp.get(), temp.bytes(), s));
    } else {
      HIP_CHECK((hipError_t)vendor::ThrustMerge<T>(da.get<T>(), na, db.get<T>(), nb, out.get<T>(), temp.get(),
                                                  temp.bytes(), s));
    }
  };
  merge_once();  // untimed: loads the code objects
  s.Sync();
  HIP_CHECK(hipMemset(temp.get(), 0, sizeof(uint32_t)));  // the k-way merge's status word
  HIP_CHECK(hipDeviceSynchronize());

  TimeDurations::Get()->Tic("local_merge_phase");
  merge_once();
  s.Sync();
  TimeDurations::Get()->Toc("local_merge_phase");
  if (leg == Leg::kMultiway) MGS_CHECK(mgs_temp_status(temp.get(), temp.bytes(), s));

  if (n) HIP_CHECK(hipMemcpy(merged.data(), out.get(), n * sizeof(T), hipMemcpyDeviceToHost));
  std::cout << n << ",\"" << name << "\"," << gpu << "," << std::fixed << std::setprecision(9)
            << TimeDurations::Get()->Seconds("local_merge_phase") << "\n";
  CheckOrExit(IsSortedEncoded(merged.data(), n) && Fingerprint(merged.data(), n) == before);
  return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  const std::vector<std::string> a(argv + 1, argv + argc);
  if (a.size() < 3 || a.size() > 4) UsageExit(kUsage);
  size_t n = 0, gpu = 0;
  Leg leg;
  if (!ParseCount(a[0], &n) || !ParseLeg(a[1], &leg) || !ParseCount(a[2], &gpu)) UsageExit(kUsage);
  const std::string type = a.size() > 3 ? a[3] : "int";
  try {
    if (type == "int") return Run<int32_t>(n, a[1], leg, (int)gpu);
    if (type == "long") return Run<int64_t>(n, a[1], leg, (int)gpu);
  } catch (const std::exception& e) {
    std::cout << "Error: " << e.what() << "\n";
    return 1;
  }
  UsageExit(kUsage);
}
