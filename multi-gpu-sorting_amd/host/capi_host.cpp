// extern "C" surface of the host layer (libmgsort_host.so) for the Python
// tests and benchmarks: the data generator and the CPU baselines.
#include <cstring>
#include <string>

#include "data_generator.h"
#include "host_api.h"

namespace {

template <typename T>
int Gen(const char* dist, uint32_t seed, void* out, size_t n) {
  mgsort::DataGenerator g(seed);
  return g.ComputeDistribution<T>(static_cast<T*>(out), n, dist) ? 0 : -1;
}

}  // namespace

extern "C" int mgh_generate(const char* dtype, const char* dist, uint32_t seed, void* out, size_t n) {
  if (dtype == nullptr || dist == nullptr || (out == nullptr && n > 0)) return -4;
  const std::string t(dtype);
  if (t == "int" || t == "i32") return Gen<int>(dist, seed, out, n);
  if (t == "long" || t == "i64") return Gen<long>(dist, seed, out, n);
  if (t == "float" || t == "f32") return Gen<float>(dist, seed, out, n);
  if (t == "double" || t == "f64") return Gen<double>(dist, seed, out, n);
  if (t == "u32") return Gen<uint32_t>(dist, seed, out, n);
  if (t == "u64") return Gen<uint64_t>(dist, seed, out, n);
  return -1;
}
