// data_transfer_benchmark: host-link transfers of one GPU, timed
// (src/data_transfer_benchmark.cu:13-241).  Same positional arguments and
// CSV line as the reference:
//   num_bytes,"gpus",num_repetitions,execution_type,"h_to_d_durations",
//   h_to_d_total_duration,"p_to_p_durations",p_to_p_total_duration,
//   "d_to_h_durations",d_to_h_total_duration,"bidirectional_durations",
//   total_bidirectional_duration
// This build measures one GPU: `gpus` must name exactly one device, and the
// peer-to-peer fields are empty and 0, as the reference prints them for one
// GPU.  Per the reference: one host-to-device copy of num_bytes, one
// device-to-host copy of num_bytes, then num_repetitions rounds of both at
// once on two streams (the bidirectional durations are per round).  Host
// memory is pinned.  Each copy is timed on the host around the copy and a
// wait for its stream.  Exits 1 if the bytes that arrived on the device
// differ from what was sent.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "argument_limits.h"
#include "bench_common.h"
#include "device_buffer.h"
#include "hip_check.h"
#include "host_vector.h"
#include "time_durations.h"

using namespace mgsort;

int main(int argc, char* argv[]) {
  const char usage[] =
      "./data_transfer_benchmark <num_bytes> <gpus> <num_repetitions> <execution_type=serial,parallel>\n"
      "  (one GPU: gpus names a single device)\n";
  const std::vector<std::string> a(argv + 1, argv + argc);
  if (a.size() != 4) UsageExit(usage);
  size_t nb = 0, reps = 0;
  std::vector<int> gpus;
  if (!ParseCount(a[0], &nb) || nb == 0 || !ArgumentLimits::ParseGpus(a[1], &gpus) || !ParseCount(a[2], &reps) ||
      reps == 0 || (a[3] != "serial" && a[3] != "parallel"))
    UsageExit(usage);
  if (gpus.size() != 1) {
    std::cout << "Error: this build measures one GPU; peer-to-peer transfers need a second device.\n";
    return 1;
  }
  const int gpu = gpus.front();
  TimeDurations* td = TimeDurations::Get();
  try {
    HostVector<uint8_t> host(2 * nb, false);
    if (!host.pinned()) throw Error("pinned host allocation failed");
    uint8_t* h = host.data();
#pragma omp parallel for
    for (size_t i = 0; i < 2 * nb; ++i) h[i] = (uint8_t)(i * 131 + 7);
    HIP_CHECK(hipSetDevice(gpu));
    HIP_CHECK(hipDeviceSynchronize());
    DeviceBuffer primary(2 * nb), secondary(2 * nb);
    HIP_CHECK(hipMemset(secondary.get(), 0, 2 * nb));
    Stream s, htod, dtoh;
    const std::string id = std::to_string(gpu);

    td->Tic("total_host_to_device_transfer");
    td->Tic("host_to_device_transfer_" + id);
    HIP_CHECK(hipMemcpyAsync(primary.get(), host.data(), nb, hipMemcpyHostToDevice, s));
    s.Sync();
    td->Toc("host_to_device_transfer_" + id);
    td->Toc("total_host_to_device_transfer");
    uint8_t first = 0, last = 0;
    HIP_CHECK(hipMemcpy(&first, primary.get(), 1, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(&last, primary.get<uint8_t>() + nb - 1, 1, hipMemcpyDeviceToHost));
    if (first != host.data()[0] || last != host.data()[nb - 1]) throw Error("host-to-device copy lost bytes");

    td->Tic("total_device_to_host_transfer");
    td->Tic("device_to_host_transfer_" + id);
    HIP_CHECK(hipMemcpyAsync(host.data(), secondary.get(), nb, hipMemcpyDeviceToHost, s));
    s.Sync();
    td->Toc("device_to_host_transfer_" + id);
    td->Toc("total_device_to_host_transfer");

    td->Tic("total_bidirectional_data_transfer");
    td->Tic("bidirectional_data_transfer_" + id);
    for (size_t r = 0; r < reps; ++r) {
      HIP_CHECK(hipMemcpyAsync(primary.get(), host.data(), nb, hipMemcpyHostToDevice, htod));
      HIP_CHECK(hipMemcpyAsync(host.data() + nb, primary.get<uint8_t>() + nb, nb, hipMemcpyDeviceToHost, dtoh));
      htod.Sync();
      dtoh.Sync();
    }
    td->Toc("bidirectional_data_transfer_" + id);
    td->Toc("total_bidirectional_data_transfer");
  } catch (const std::exception& e) {
    std::cout << "Error: " << e.what() << "\n";
    return 1;
  }
  const std::string id = std::to_string(gpu);
  const double r = (double)reps;
  std::cout << std::fixed << std::setprecision(9) << nb << ",\"" << gpu << "\"," << reps << "," << a[3] << ",\""
            << td->Seconds("host_to_device_transfer_" + id) << "\"," << td->Seconds("total_host_to_device_transfer")
            << ",\"\"," << td->Seconds("total_peer_to_peer_transfer") << ",\""
            << td->Seconds("device_to_host_transfer_" + id) << "\"," << td->Seconds("total_device_to_host_transfer")
            << ",\"" << td->Seconds("bidirectional_data_transfer_" + id) / r << "\","
            << td->Seconds("total_bidirectional_data_transfer") / r << "\n";
  return 0;
}
