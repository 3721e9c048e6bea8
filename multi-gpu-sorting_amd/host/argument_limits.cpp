// Command-line surface of sort_benchmark (see argument_limits.h).
#include "argument_limits.h"

#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <cctype>
#include <limits>
#include <numeric>
#include <set>
#include <sstream>

#include "data_generator.h"

namespace mgsort {

namespace {

bool AllDigits(const std::string& s) {
  return !s.empty() && std::all_of(s.begin(), s.end(), [](unsigned char c) { return std::isdigit(c); });
}

bool ParseUnsigned(const std::string& s, unsigned long long max, unsigned long long* v) {
  if (!AllDigits(s) || s.size() > 20) return false;
  try {
    *v = std::stoull(s);
  } catch (...) {
    return false;
  }
  return *v <= max;
}

bool IsPowerOfTwo(size_t x) { return x != 0 && (x & (x - 1)) == 0; }

template <typename T>
std::string Join(const std::vector<T>& v) {
  std::stringstream s;
  s << "{";
  for (size_t i = 0; i < v.size(); ++i) s << (i ? ", " : "") << v[i];
  s << "}";
  return s.str();
}

}  // namespace

const std::vector<std::string>& ArgumentLimits::ValidAlgorithms() {
  static const std::vector<std::string> v = {"tanasic", "gowanlock", "thrust", "boost", "gnu_parallel",
                                             "std",     "paradis",   "hip"};
  return v;
}

const std::vector<std::string>& ArgumentLimits::ValidDataTypes() {
  static const std::vector<std::string> v = {"int", "long", "float", "double", "u32", "u64"};
  return v;
}

std::vector<std::string> ArgumentLimits::ValidDistributionTypes() {
  std::vector<std::string> v = DataGenerator::ReferenceDistributions();
  const auto& a = DataGenerator::AddedDistributions();
  v.insert(v.end(), a.begin(), a.end());
  return v;
}

const std::vector<int>& ArgumentLimits::ValidGpus() {
  static const std::vector<int> v = [] {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
    std::vector<int> g(count);
    std::iota(g.begin(), g.end(), 0);
    return g;
  }();
  return v;
}

std::vector<int> ArgumentLimits::GetDefaultGpus() {
  const auto& v = ValidGpus();
  size_t n = 1;
  while (n * 2 <= v.size()) n *= 2;
  if (v.empty()) return {0};
  return {v.begin(), v.begin() + n};
}

int ArgumentLimits::GetDefaultNumThreads() { return omp_get_num_procs(); }

bool ArgumentLimits::ParseNumElements(const std::string& s, size_t* v) {
  unsigned long long x;
  if (!ParseUnsigned(s, std::numeric_limits<size_t>::max(), &x)) return false;
  *v = x;
  return true;
}

bool ArgumentLimits::ParseAlgorithm(const std::string& s, std::string* v) {
  const auto& a = ValidAlgorithms();
  if (std::find(a.begin(), a.end(), s) == a.end()) return false;
  *v = s;
  return true;
}

bool ArgumentLimits::ParseGpus(const std::string& in, std::vector<int>* gpus) {
  std::string s;
  for (char c : in)
    if (!std::isspace((unsigned char)c)) s.push_back(c);
  if (s.empty()) return false;
  if (s.front() == '{' || s.back() == '}') {
    if ((s.front() == '{') != (s.back() == '}') || s.size() < 2) return false;
    s = s.substr(1, s.size() - 2);
  }
  if (s.empty() || s.front() == ',' || s.back() == ',' || s.find(",,") != std::string::npos) return false;
  if (!std::all_of(s.begin(), s.end(), [](unsigned char c) { return std::isdigit(c) || c == ','; })) return false;
  std::vector<int> out;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, ',')) {
    unsigned long long x;
    if (!ParseUnsigned(item, 1u << 20, &x)) return false;
    out.push_back((int)x);
  }
  if (!IsPowerOfTwo(out.size())) return false;
  if (std::set<int>(out.begin(), out.end()).size() != out.size()) return false;
  const auto& valid = ValidGpus();
  for (int g : out)
    if (std::find(valid.begin(), valid.end(), g) == valid.end()) return false;
  *gpus = out;
  return true;
}

bool ArgumentLimits::ParseDataType(const std::string& s, std::string* v) {
  const auto& a = ValidDataTypes();
  if (std::find(a.begin(), a.end(), s) == a.end()) return false;
  *v = s;
  return true;
}

bool ArgumentLimits::ParseDistributionType(const std::string& s, std::string* v) {
  if (!DataGenerator::IsValidDistribution(s)) return false;
  *v = s;
  return true;
}

bool ArgumentLimits::ParseDistributionSeed(const std::string& s, uint32_t* v) {
  unsigned long long x;
  if (!ParseUnsigned(s, std::numeric_limits<uint32_t>::max(), &x)) return false;
  *v = (uint32_t)x;
  return true;
}

bool ArgumentLimits::ParseNumThreads(const std::string& s, int* v) {
  unsigned long long x;
  if (!ParseUnsigned(s, (unsigned long long)GetDefaultNumThreads(), &x) || x < 1) return false;
  *v = (int)x;
  return true;
}

bool ArgumentLimits::ParseChunkSize(const std::string& s, size_t* v) {
  unsigned long long x;
  if (!ParseUnsigned(s, std::numeric_limits<size_t>::max(), &x)) return false;
  *v = x;
  return true;
}

bool ArgumentLimits::ParseMergeGroupSize(const std::string& s, size_t* v) { return ParseChunkSize(s, v); }

bool ArgumentLimits::ParseNumBuffers(const std::string& s, size_t* v) {
  unsigned long long x;
  if (!ParseUnsigned(s, 3, &x) || x < 2) return false;
  *v = x;
  return true;
}

bool ArgumentLimits::ParseValues(const std::string& s, std::string* v) {
  if (s != "none" && s != "u32" && s != "u64") return false;
  *v = s;
  return true;
}

bool ArgumentLimits::ParsePartitions(const std::string& s, size_t* v) {
  unsigned long long x;
  if (!ParseUnsigned(s, 1u << 10, &x)) return false;
  if (x != 0 && !IsPowerOfTwo(x)) return false;
  *v = x;
  return true;
}

bool ArgumentLimits::ParseMergeBackend(const std::string& s, std::string* v) {
  if (s != "cpu" && s != "gpu") return false;
  *v = s;
  return true;
}

std::string ArgumentLimits::Usage() {
  std::stringstream u;
  u << "Usage: ./sort_benchmark <num_elements> [options]\n\n"
    << "<num_elements>\t\t\t\t\tThe number of elements [0, " << std::numeric_limits<size_t>::max() << "]\n\n"
    << "--algorithm\t\t<algorithm>\t\tThe algorithm " << Join(ValidAlgorithms()) << " (= " << GetDefaultAlgorithm()
    << ")\n"
    << "--gpus\t\t\t<gpus>\t\t\tThe visible GPUs " << Join(ValidGpus()) << " (= " << Join(GetDefaultGpus()) << ")\n"
    << "--num_threads\t\t<num_threads>\t\tThe number of threads [1, " << GetDefaultNumThreads()
    << "] (= " << GetDefaultNumThreads() << ")\n"
    << "--data_type\t\t<data_type>\t\tThe data type " << Join(ValidDataTypes()) << " (= " << GetDefaultDataType()
    << ")\n"
    << "--distribution_type\t<distribution_type>\tThe distribution type " << Join(ValidDistributionTypes())
    << " (= " << GetDefaultDistributionType() << ")\n"
    << "--distribution_seed\t<distribution_seed>\tThe distribution seed [0, 4294967295] (= "
    << GetDefaultDistributionSeed() << ")\n"
    << "--values\t\t<values>\t\tPayload per key {none, u32, u64} (= none; algorithms hip, thrust)\n"
    << "\n[--algorithm tanasic]\n"
    << "--partitions\t\t<partitions>\t\tVirtual partitions on one GPU, a power of two (= one per GPU)\n"
    << "\n[--algorithm gowanlock]\n"
    << "--chunk_size\t\t<chunk_size>\t\tThe chunk_size [0, max] (= auto)\n"
    << "--merge_group_size\t<merge_group_size>\tThe merge group size [0, max] (= " << GetDefaultMergeGroupSize()
    << ")\n"
    << "--num_buffers\t\t<num_buffers>\t\tThe number of device buffers [2, 3] (= " << GetDefaultNumBuffers()
    << ")\n"
    << "--merge_backend\t\t<merge_backend>\t\tFinal merge {cpu, gpu} (= cpu)\n"
    << "--block_size\t\t<block_size>\t\tBytes per in-place transfer block of --num_buffers 3 (= chunk/100)\n";
  return u.str();
}

}  // namespace mgsort
