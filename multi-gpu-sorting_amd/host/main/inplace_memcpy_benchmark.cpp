// inplace_memcpy_benchmark: the in-place exchange through one device buffer
// (the HET sort's 3-buffer schedule), timed
// (src/inplace_memcpy_benchmark.cu:15-79).  Same positional arguments and CSV
// line:
//   num_bytes,gpu_id,block_size,inplace_data_transfer_duration
// The device buffer starts as num_bytes of 42 and the pinned host buffer as
// 2 num_bytes of 33; the exchange sends the device bytes to the host
// buffer's second half while the first half refills the device buffer,
// block_size bytes at a time (EnqueueInplaceTransfer: ordered by events, not
// by host waits as in the reference).  block_size may be given as a decimal
// fraction, as the reference's experiment grid writes it (b / i in Python);
// the fraction is dropped, as std::stoull does.  Exits 1 unless every byte
// arrived ("ERROR", as the reference prints).
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
#include "device_buffer.h"
#include "hip_check.h"
#include "host_vector.h"
#include "inplace_transfer.h"
#include "time_durations.h"

using namespace mgsort;

namespace {

// "123" or "123.45" -> 123
bool ParseTruncated(const std::string& s, size_t* v) {
  const size_t dot = s.find('.');
  if (dot == std::string::npos) return ParseCount(s, v);
  for (size_t i = dot + 1; i < s.size(); ++i)
    if (s[i] < '0' || s[i] > '9') return false;
  return ParseCount(s.substr(0, dot), v);
}

}  // namespace

int main(int argc, char* argv[]) {
  const char usage[] = "./inplace_memcpy_benchmark <num_bytes> <gpu_id> <block_size>\n";
  const std::vector<std::string> a(argv + 1, argv + argc);
  if (a.size() != 3) UsageExit(usage);
  size_t nb = 0, block = 0;
  std::vector<int> gpus;
  if (!ParseCount(a[0], &nb) || nb == 0 || !ArgumentLimits::ParseGpus(a[1], &gpus) || gpus.empty() ||
      !ParseTruncated(a[2], &block) || block == 0)
    UsageExit(usage);
  block = std::min(block, nb);
  const int gpu = gpus.front();
  try {
    HostVector<uint8_t> host(2 * nb, false);
    if (!host.pinned()) throw Error("pinned host allocation failed");
    std::fill(host.data(), host.data() + 2 * nb, (uint8_t)33);
    HIP_CHECK(hipSetDevice(gpu));
    HIP_CHECK(hipDeviceSynchronize());
    DeviceBuffer dev(nb);
    HIP_CHECK(hipMemset(dev.get(), 42, nb));
    Stream h2d, d2h;
    std::vector<std::unique_ptr<Event>> block_out;
    for (size_t k = 0; k < (nb + block - 1) / block; ++k) block_out.emplace_back(new Event());
    HIP_CHECK(hipDeviceSynchronize());

    TimeDurations::Get()->Tic("data_transfer");
    EnqueueInplaceTransfer(dev.get<char>(), reinterpret_cast<const char*>(host.data()), nb,
                           reinterpret_cast<char*>(host.data()) + nb, nb, block, h2d, d2h, block_out);
    h2d.Sync();
    d2h.Sync();
    TimeDurations::Get()->Toc("data_transfer");

    std::vector<uint8_t> back(nb);
    HIP_CHECK(hipMemcpy(back.data(), dev.get(), nb, hipMemcpyDeviceToHost));
    const bool ok = std::all_of(host.data() + nb, host.data() + 2 * nb, [](uint8_t x) { return x == 42; }) &&
                    std::all_of(back.begin(), back.end(), [](uint8_t x) { return x == 33; });
    if (!ok) {
      std::cout << "ERROR" << std::endl;
      return 1;
    }
  } catch (const std::exception& e) {
    std::cout << "Error: " << e.what() << "\n";
    return 1;
  }
  std::cout << std::fixed << std::setprecision(9) << nb << "," << gpu << "," << block << ","
            << TimeDurations::Get()->Seconds("data_transfer") << std::endl;
  return 0;
}
