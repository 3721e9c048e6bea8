// serve_benchmark: the serving layer (SortServer) under a Poisson stream of
// mixed sort requests, batched (continuous batching into segmented launches)
// against the plain one-request-per-launch path.
//
// Request mix (seeded mt19937_64): n = round(2^(lo + (hi-lo) U)) keys;
// 50% u32 keys, 30% u64 keys, 20% u32 keys with u32 payloads (payload = input
// index, so stability is observable); each over the reference generator's
// uniform distribution or sorted-then-1%-swapped ("nearly-sorted-swaps").
//
// The expected output of every request is computed on the host first
// (std::sort of (key, input index) pairs: the stable order by key, which for
// keys alone is also the exact multiset check).  Then:
//   ceiling   each request alone on the device (device data, plain path, one
//             at a time): total keys over the sum of those times, the rate
//             the requests would reach one after another at their sizes'
//             best single-request rate
//   sweep     for every data mode (host: arrays in pinned host memory,
//             copied in and out; device: arrays already in the server's HBM
//             pool, sorted in place) and every load (requests/s as a
//             multiple of the plain path's saturated rate, "inf" = all
//             submitted at once), one JSON line per mode with throughput
//             (GKeys/s), its fraction of the ceiling, p50/p99/mean latency
//             (submit -> completion event) and whether every output equals
//             the host-sorted one.
#include <hip/hip_runtime.h>
#include <omp.h>

#include <parallel/algorithm>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "data_generator.h"
#include "hip_check.h"
#include "host_vector.h"
#include "key_traits.h"
#include "sort_server.h"

using namespace mgsort;

namespace {

struct Spec {
  mgs_key kt;
  mgs_val vt;
  size_t n;
  std::string dist;
  uint32_t seed;
};

struct Data {
  HostVector<uint8_t> keys, vals, out_keys, out_vals, exp_keys, exp_vals;
};

size_t KB(mgs_key kt) { return kt == MGS_KEY_U64 ? 8 : 4; }
size_t VB(mgs_val vt) { return vt == MGS_VAL_U32 ? 4 : 0; }

std::vector<Spec> MakeSpecs(int count, double lo, double hi, uint64_t seed) {
  std::mt19937_64 g(seed);
  std::vector<Spec> v;
  for (int i = 0; i < count; ++i) {
    const double u = (g() >> 11) * (1.0 / 9007199254740992.0);
    Spec s;
    s.n = (size_t)std::llround(std::pow(2.0, lo + (hi - lo) * u));
    const double t = (g() >> 11) * (1.0 / 9007199254740992.0);
    if (t < 0.5) {
      s.kt = MGS_KEY_U32;
      s.vt = MGS_VAL_NONE;
    } else if (t < 0.8) {
      s.kt = MGS_KEY_U64;
      s.vt = MGS_VAL_NONE;
    } else {
      s.kt = MGS_KEY_U32;
      s.vt = MGS_VAL_U32;
    }
    s.dist = (g() & 1) ? "nearly-sorted-swaps" : "uniform";
    s.seed = (uint32_t)(g() & 0x7FFFFFFF);
    v.push_back(s);
  }
  return v;
}

void Generate(const Spec& s, Data& d) {
  d.keys.Resize(s.n * KB(s.kt), false);
  DataGenerator gen(s.seed);
  bool ok = s.kt == MGS_KEY_U64 ? gen.ComputeDistribution<uint64_t>(reinterpret_cast<uint64_t*>(d.keys.data()), s.n, s.dist)
                                : gen.ComputeDistribution<uint32_t>(reinterpret_cast<uint32_t*>(d.keys.data()), s.n, s.dist);
  if (!ok) throw Error("unknown distribution " + s.dist);
  if (s.vt == MGS_VAL_U32) {
    d.vals.Resize(s.n * 4, false);
    uint32_t* v = reinterpret_cast<uint32_t*>(d.vals.data());
#pragma omp parallel for
    for (size_t i = 0; i < s.n; ++i) v[i] = (uint32_t)i;
  }
  d.out_keys.Resize(s.n * KB(s.kt), false);
  d.exp_keys.Resize(s.n * KB(s.kt), false);
  if (s.vt != MGS_VAL_NONE) {
    d.out_vals.Resize(s.n * VB(s.vt), false);
    d.exp_vals.Resize(s.n * VB(s.vt), false);
  }
}

// Expected output: the (key, input index) pairs sorted, i.e. the stable
// order by key (unsigned keys: the generator's u32 / u64).
template <typename K>
void HostSort(const Spec& s, Data& d) {
  const K* k = reinterpret_cast<const K*>(d.keys.data());
  std::vector<std::pair<K, uint32_t>> p(s.n);
  for (size_t i = 0; i < s.n; ++i) p[i] = {k[i], (uint32_t)i};
  if (s.n >= (size_t(1) << 20))
    __gnu_parallel::sort(p.begin(), p.end());
  else
    std::sort(p.begin(), p.end());
  K* ek = reinterpret_cast<K*>(d.exp_keys.data());
  uint32_t* ev = s.vt ? reinterpret_cast<uint32_t*>(d.exp_vals.data()) : nullptr;
  for (size_t i = 0; i < s.n; ++i) {
    ek[i] = p[i].first;
    if (ev) ev[i] = reinterpret_cast<const uint32_t*>(d.vals.data())[p[i].second];
  }
}

void Expect(const Spec& s, Data& d) {
  if (s.kt == MGS_KEY_U64)
    HostSort<uint64_t>(s, d);
  else
    HostSort<uint32_t>(s, d);
}

struct RunResult {
  double wall = 0, p50 = 0, p99 = 0, mean = 0, sum_latency = 0;
  uint64_t keys = 0, batches = 0, failed = 0;
  bool verified = true;
};

struct Options {
  size_t budget = size_t(64) << 30;
  size_t solo_min_keys = size_t(1) << 21;
  int sort_streams = 2;
  std::string seg_tuning;
};

double Percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const size_t i = std::min(v.size() - 1, (size_t)std::ceil(q * v.size()) - 1);
  return v[i];
}

// one_at_a_time: each request is submitted after the previous one completed
// (each sorted alone).
RunResult Run(const std::vector<Spec>& specs, std::vector<Data>& data, bool batched, int inflight, bool host_data,
              double lambda, const Options& opt, uint64_t seed, bool one_at_a_time) {
  ServerConfig cfg;
  cfg.batched = batched;
  cfg.max_inflight = inflight;
  cfg.host_data = host_data;
  cfg.hbm_budget = opt.budget;
  cfg.solo_min_keys = opt.solo_min_keys;
  cfg.sort_streams = batched ? opt.sort_streams : 1;
  cfg.seg_tuning_path = opt.seg_tuning;
  SortServer server(cfg);
  const size_t R = specs.size();
  std::vector<SortRequest> reqs(R);
  std::vector<void*> slots(R, nullptr);
  for (size_t i = 0; i < R; ++i) {
    const Spec& s = specs[i];
    SortRequest& r = reqs[i];
    r.kt = s.kt;
    r.vt = s.vt;
    r.n = s.n;
    r.keys_in = data[i].keys.data();
    r.vals_in = s.vt ? data[i].vals.data() : nullptr;
    r.keys_out = data[i].out_keys.data();
    r.vals_out = s.vt ? data[i].out_vals.data() : nullptr;
    if (!host_data) {  // the client's data already sits in the pool
      slots[i] = server.AllocSlot(s.kt, s.vt, s.n);
      if (!slots[i]) throw Error("pool too small for the device-mode request set");
      HIP_CHECK(hipMemcpy(slots[i], r.keys_in, s.n * KB(s.kt), hipMemcpyHostToDevice));
      if (s.vt)
        HIP_CHECK(hipMemcpy(SortServer::SlotVals(slots[i], s.kt, s.n), r.vals_in, s.n * VB(s.vt),
                            hipMemcpyHostToDevice));
      r.slot = slots[i];
    }
  }
  // Poisson arrival offsets
  std::vector<double> at(R, 0.0);
  if (std::isfinite(lambda)) {
    std::mt19937_64 g(seed);
    std::exponential_distribution<double> gap(lambda);
    double t = 0;
    for (size_t i = 0; i < R; ++i) {
      at[i] = t;
      t += gap(g);
    }
  }
  const double t0 = SortServer::Now();
  for (size_t i = 0; i < R; ++i) {
    const double wait = t0 + at[i] - SortServer::Now();
    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
    server.Submit(&reqs[i]);
    if (one_at_a_time) server.Drain();
  }
  server.Drain();
  RunResult res;
  std::vector<double> lat;
  double first = 1e300, last = 0;
  for (auto& r : reqs) {
    lat.push_back(r.t_done - r.t_submit);
    first = std::min(first, r.t_submit);
    last = std::max(last, r.t_done);
    if (r.status == MGS_OK) res.keys += r.n;
  }
  const ServerStats st = server.Stats();
  res.batches = st.batches;
  res.failed = st.failed;
  res.wall = last - first;
  res.p50 = Percentile(lat, 0.50) * 1e3;
  res.p99 = Percentile(lat, 0.99) * 1e3;
  double sum = 0;
  for (double x : lat) sum += x;
  res.sum_latency = sum;
  res.mean = sum / std::max<size_t>(1, lat.size()) * 1e3;
  // results: device mode reads them back from the pool
  for (size_t i = 0; i < R; ++i) {
    const Spec& s = specs[i];
    if (!host_data) {
      HIP_CHECK(hipMemcpy(data[i].out_keys.data(), slots[i], s.n * KB(s.kt), hipMemcpyDeviceToHost));
      if (s.vt)
        HIP_CHECK(hipMemcpy(data[i].out_vals.data(), SortServer::SlotVals(slots[i], s.kt, s.n), s.n * VB(s.vt),
                            hipMemcpyDeviceToHost));
      server.FreeSlot(slots[i]);
    }
    const size_t kb = s.n * KB(s.kt), vb = s.n * VB(s.vt);
    res.verified = res.verified && reqs[i].status == MGS_OK &&
                   std::memcmp(data[i].exp_keys.data(), data[i].out_keys.data(), kb) == 0 &&
                   (vb == 0 || std::memcmp(data[i].exp_vals.data(), data[i].out_vals.data(), vb) == 0);
  }
  return res;
}

// The package's tuning table beside this binary's directory
// (bin/../mgsort/seg_tuning.json), if present.
std::string DefaultTuningPath() {
  char buf[4096];
  const ssize_t len = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
  if (len <= 0) return "";
  buf[len] = 0;
  std::string dir(buf);
  dir = dir.substr(0, dir.rfind('/'));
  const std::string p = dir + "/../mgsort/seg_tuning.json";
  return std::ifstream(p).good() ? p : "";
}

bool Arg(int argc, char** argv, const char* name, std::string* v) {
  for (int i = 1; i + 1 < argc; ++i)
    if (std::strcmp(argv[i], name) == 0) {
      *v = argv[i + 1];
      return true;
    }
  return false;
}

std::vector<std::string> Split(const std::string& s) {
  std::vector<std::string> out;
  size_t a = 0;
  while (a <= s.size()) {
    const size_t b = s.find(',', a);
    out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return out;
}

}  // namespace

int main(int argc, char** argv) {
  std::string v;
  int requests = 192;
  double lo = 10, hi = 26, budget_gib = 64;
  uint64_t seed = 7;
  std::string data_modes = "device,host", loads = "inf,2,1,0.5";
  Options opt;
  opt.seg_tuning = DefaultTuningPath();
  if (Arg(argc, argv, "--requests", &v)) requests = std::stoi(v);
  // requests of at least 2^solo_min_log2 keys are sorted alone (-1: never)
  if (Arg(argc, argv, "--solo_min_log2", &v)) {
    const int e = std::stoi(v);
    opt.solo_min_keys = e < 0 ? 0 : size_t(1) << e;
  }
  if (Arg(argc, argv, "--sort_streams", &v)) opt.sort_streams = std::stoi(v);
  if (Arg(argc, argv, "--seg_tuning", &v)) opt.seg_tuning = v == "none" ? "" : v;
  if (Arg(argc, argv, "--min_log2", &v)) lo = std::stod(v);
  if (Arg(argc, argv, "--max_log2", &v)) hi = std::stod(v);
  if (Arg(argc, argv, "--budget_gib", &v)) budget_gib = std::stod(v);
  if (Arg(argc, argv, "--seed", &v)) seed = std::stoull(v);
  if (Arg(argc, argv, "--data", &v)) data_modes = v;
  if (Arg(argc, argv, "--loads", &v)) loads = v;
  opt.budget = (size_t)(budget_gib * (1ull << 30));

  try {
    HIP_CHECK(hipSetDevice(0));
    const std::vector<Spec> specs = MakeSpecs(requests, lo, hi, seed);
    std::vector<Data> data(specs.size());
    uint64_t total = 0;
    for (size_t i = 0; i < specs.size(); ++i) {
      Generate(specs[i], data[i]);
      total += specs[i].n;
    }
    std::printf("{\"event\": \"generated\", \"requests\": %d, \"keys\": %llu, \"threads\": %d, "
                "\"solo_min_keys\": %zu, \"sort_streams\": %d, \"seg_tuning\": \"%s\"}\n", requests,
                (unsigned long long)total, omp_get_max_threads(), opt.solo_min_keys, opt.sort_streams,
                opt.seg_tuning.c_str());
    std::fflush(stdout);
    {
      const double t0 = SortServer::Now();
      for (size_t i = 0; i < specs.size(); ++i) Expect(specs[i], data[i]);
      std::printf("{\"event\": \"expected\", \"what\": \"host sort of (key, index) pairs\", \"s\": %.3f}\n",
                  SortServer::Now() - t0);
      std::fflush(stdout);
    }
    // each request alone on the device, one at a time through the plain path
    double ceiling = 0;
    {
      const RunResult r = Run(specs, data, false, 1, false, INFINITY, opt, seed, true);
      ceiling = total / r.sum_latency / 1e9;
      std::printf("{\"event\": \"ceiling\", \"what\": \"each request alone on the device (data in HBM, mgs_sort, one "
                  "at a time): total keys / sum of their latencies\", \"alone_sum_s\": %.6f, \"gkeys_s\": %.4f, "
                  "\"matches_host_sort\": %s}\n",
                  r.sum_latency, ceiling, r.verified ? "true" : "false");
      std::fflush(stdout);
      if (!r.verified) return 1;
    }
    bool all_ok = true;
    for (const std::string& dm : Split(data_modes)) {
      const bool host = dm == "host";
      // saturated plain rate fixes the request rates of the sweep
      const RunResult sat = Run(specs, data, false, 1, host, INFINITY, opt, seed, false);
      const double plain_rps = specs.size() / sat.wall;
      struct Mode {
        const char* name;
        bool batched;
        int inflight;
      };
      const Mode modes[] = {{"plain", false, 1}, {"plain_overlap", false, 3}, {"batched", true, 3}};
      for (const std::string& ld : Split(loads)) {
        const bool inf = ld == "inf";
        const double lambda = inf ? INFINITY : std::stod(ld) * plain_rps;
        for (const Mode& m : modes) {
          const RunResult r = (inf && !m.batched && m.inflight == 1)
                                  ? sat
                                  : Run(specs, data, m.batched, m.inflight, host, lambda, opt, seed + 1, false);
          all_ok = all_ok && r.verified;
          const double gk = r.keys / r.wall / 1e9;
          std::printf(
              "{\"data\": \"%s\", \"mode\": \"%s\", \"load\": \"%s\", \"lambda_rps\": %s, \"requests\": %zu, "
              "\"keys\": %llu, \"wall_s\": %.6f, \"gkeys_s\": %.4f, \"frac_of_ceiling\": %.4f, \"p50_ms\": %.3f, "
              "\"p99_ms\": %.3f, \"mean_ms\": %.3f, \"batches\": %llu, \"failed\": %llu, \"matches_host_sort\": %s}\n",
              dm.c_str(), m.name, ld.c_str(), inf ? "null" : std::to_string(lambda).c_str(), specs.size(),
              (unsigned long long)r.keys, r.wall, gk, gk / ceiling, r.p50, r.p99, r.mean,
              (unsigned long long)r.batches, (unsigned long long)r.failed, r.verified ? "true" : "false");
          std::fflush(stdout);
        }
      }
    }
    return all_ok ? 0 : 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "serve_benchmark: %s\n", e.what());
    return 2;
  }
}
