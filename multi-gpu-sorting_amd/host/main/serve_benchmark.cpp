// serve_benchmark: the serving layer (SortServer) under a Poisson stream of
// mixed sort requests, batched (continuous batching into segmented launches)
// against the plain one-request-per-launch path.
//
// Request mix (seeded mt19937_64): n = round(2^(lo + (hi-lo) U)) keys;
// 50% u32 keys, 30% u64 keys, 20% u32 keys with u32 payloads (payload = input
// index, so stability is observable); each over the reference generator's
// uniform distribution or sorted-then-1%-swapped ("nearly-sorted-swaps").
//
// For every data mode (host: arrays in pinned host memory, copied in and out;
// device: arrays already in the server's HBM pool, sorted in place) and every
// load (requests/s as a multiple of the plain path's saturated rate, "inf" =
// all submitted at once) it prints one JSON line with throughput (GKeys/s),
// p50/p99/mean latency (submit -> completion event) and whether every output
// equals the output of the same request sorted alone.
#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
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

// Sorted by key, and for pairs ties in payload (= input index) order.
template <typename K>
bool CheckSorted(const K* k, const uint32_t* v, size_t n) {
  bool ok = true;
#pragma omp parallel for reduction(&& : ok)
  for (size_t i = 1; i < n; ++i) {
    const bool in_order = k[i - 1] < k[i] || (k[i - 1] == k[i] && (v == nullptr || v[i - 1] < v[i]));
    ok = ok && in_order;
  }
  return ok;
}

struct RunResult {
  double wall = 0, p50 = 0, p99 = 0, mean = 0;
  uint64_t keys = 0, batches = 0, failed = 0;
  bool verified = true;
};

double Percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const size_t i = std::min(v.size() - 1, (size_t)std::ceil(q * v.size()) - 1);
  return v[i];
}

// record_expected: one request at a time (each sorted alone), keeping its
// output as the expected output of later runs.
RunResult Run(const std::vector<Spec>& specs, std::vector<Data>& data, bool batched, int inflight, bool host_data,
              double lambda, size_t budget, uint64_t seed, bool record_expected) {
  ServerConfig cfg;
  cfg.batched = batched;
  cfg.max_inflight = inflight;
  cfg.host_data = host_data;
  cfg.hbm_budget = budget;
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
    if (record_expected) server.Drain();
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
    if (record_expected) {
      std::memcpy(data[i].exp_keys.data(), data[i].out_keys.data(), kb);
      if (vb) std::memcpy(data[i].exp_vals.data(), data[i].out_vals.data(), vb);
      const uint32_t* v = s.vt ? reinterpret_cast<const uint32_t*>(data[i].out_vals.data()) : nullptr;
      const bool sorted = s.kt == MGS_KEY_U64
                              ? CheckSorted(reinterpret_cast<const uint64_t*>(data[i].out_keys.data()), v, s.n)
                              : CheckSorted(reinterpret_cast<const uint32_t*>(data[i].out_keys.data()), v, s.n);
      res.verified = res.verified && sorted && reqs[i].status == MGS_OK;
    } else {
      res.verified = res.verified && reqs[i].status == MGS_OK &&
                     std::memcmp(data[i].exp_keys.data(), data[i].out_keys.data(), kb) == 0 &&
                     (vb == 0 || std::memcmp(data[i].exp_vals.data(), data[i].out_vals.data(), vb) == 0);
    }
  }
  return res;
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
  if (Arg(argc, argv, "--requests", &v)) requests = std::stoi(v);
  if (Arg(argc, argv, "--min_log2", &v)) lo = std::stod(v);
  if (Arg(argc, argv, "--max_log2", &v)) hi = std::stod(v);
  if (Arg(argc, argv, "--budget_gib", &v)) budget_gib = std::stod(v);
  if (Arg(argc, argv, "--seed", &v)) seed = std::stoull(v);
  if (Arg(argc, argv, "--data", &v)) data_modes = v;
  if (Arg(argc, argv, "--loads", &v)) loads = v;
  const size_t budget = (size_t)(budget_gib * (1ull << 30));

  try {
    HIP_CHECK(hipSetDevice(0));
    const std::vector<Spec> specs = MakeSpecs(requests, lo, hi, seed);
    std::vector<Data> data(specs.size());
    uint64_t total = 0;
    for (size_t i = 0; i < specs.size(); ++i) {
      Generate(specs[i], data[i]);
      total += specs[i].n;
    }
    std::printf("{\"event\": \"generated\", \"requests\": %d, \"keys\": %llu, \"threads\": %d}\n", requests,
                (unsigned long long)total, omp_get_max_threads());
    std::fflush(stdout);
    // each request alone: one at a time through the plain path (the outputs
    // every later run must reproduce)
    {
      const RunResult r = Run(specs, data, false, 1, true, INFINITY, budget, seed, true);
      std::printf("{\"event\": \"alone\", \"sorted_and_stable\": %s}\n", r.verified ? "true" : "false");
      std::fflush(stdout);
      if (!r.verified) return 1;
    }
    bool all_ok = true;
    for (const std::string& dm : Split(data_modes)) {
      const bool host = dm == "host";
      // saturated plain rate fixes the request rates of the sweep
      const RunResult sat = Run(specs, data, false, 1, host, INFINITY, budget, seed, false);
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
                                  : Run(specs, data, m.batched, m.inflight, host, lambda, budget, seed + 1, false);
          all_ok = all_ok && r.verified;
          std::printf(
              "{\"data\": \"%s\", \"mode\": \"%s\", \"load\": \"%s\", \"lambda_rps\": %s, \"requests\": %zu, "
              "\"keys\": %llu, \"wall_s\": %.6f, \"gkeys_s\": %.4f, \"p50_ms\": %.3f, \"p99_ms\": %.3f, "
              "\"mean_ms\": %.3f, \"batches\": %llu, \"failed\": %llu, \"same_as_alone\": %s}\n",
              dm.c_str(), m.name, ld.c_str(), inf ? "null" : std::to_string(lambda).c_str(), specs.size(),
              (unsigned long long)r.keys, r.wall, r.keys / r.wall / 1e9, r.p50, r.p99, r.mean,
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
