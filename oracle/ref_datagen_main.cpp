// Driver around the reference's own src/data_generator.cuh (compiled from
// where it lies by oracle/Makefile into oracle/_ref/; test infrastructure).
// Usage: ref_datagen <dtype> <distribution> <seed> <n> <out.bin>
// Writes the n generated values as raw bytes.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "data_generator.cuh"

template <typename T>
int Run(const char* dist, uint32_t seed, size_t n, const char* path) {
  std::vector<T> v(n);  // value-initialised, like the reference's HostVector
  DataGenerator g(seed);
  g.ComputeDistribution<T>(v.data(), n, dist);
  FILE* f = std::fopen(path, "wb");
  if (!f) return 2;
  if (n) std::fwrite(v.data(), sizeof(T), n, f);
  std::fclose(f);
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s dtype dist seed n out.bin\n", argv[0]);
    return 1;
  }
  const std::string t = argv[1];
  const uint32_t seed = (uint32_t)std::strtoul(argv[3], nullptr, 10);
  const size_t n = (size_t)std::strtoull(argv[4], nullptr, 10);
  if (t == "int") return Run<int>(argv[2], seed, n, argv[5]);
  if (t == "long") return Run<long>(argv[2], seed, n, argv[5]);
  if (t == "float") return Run<float>(argv[2], seed, n, argv[5]);
  if (t == "double") return Run<double>(argv[2], seed, n, argv[5]);
  if (t == "u32") return Run<uint32_t>(argv[2], seed, n, argv[5]);
  if (t == "u64") return Run<uint64_t>(argv[2], seed, n, argv[5]);
  return 1;
}
