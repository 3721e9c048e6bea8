// Helpers shared by the benchmark programs: output checks and argument
// parsing.  The reference's programs only check std::is_sorted (and, on the
// device, the maximum); these check the key order and that the output is a
// permutation of the input, and the programs exit 1 when a check fails.
#pragma once

#include <omp.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "key_traits.h"

namespace mgsort {

// Order-independent fingerprint of a multiset of keys (sum and xor of a
// mixed image of every key).
template <typename T>
std::pair<uint64_t, uint64_t> Fingerprint(const T* a, size_t n) {
  uint64_t s = 0, x = 0;
#pragma omp parallel for reduction(+ : s) reduction(^ : x)
  for (size_t i = 0; i < n; ++i) {
    uint64_t v = (uint64_t)Encode(a[i]);
    v ^= v >> 31;
    v *= 0x9E3779B97F4A7C15ull;
    v ^= v >> 29;
    s += v;
    x ^= v;
  }
  return {s, x};
}

template <typename T>
bool IsSortedEncoded(const T* a, size_t n) {
  bool ok = true;
#pragma omp parallel for reduction(&& : ok)
  for (size_t i = 1; i < n; ++i) ok = ok && !(Encode(a[i]) < Encode(a[i - 1]));
  return ok;
}

// Strict unsigned decimal (the reference parses with std::stoull, which
// also takes signs and trailing text; these programs reject them).
inline bool ParseCount(const std::string& s, size_t* v) {
  if (s.empty() || s.size() > 20) return false;
  unsigned long long x = 0;
  for (char c : s) {
    if (c < '0' || c > '9') return false;
    const unsigned long long d = (unsigned long long)(c - '0');
    if (x > (~0ull - d) / 10) return false;
    x = x * 10 + d;
  }
  *v = (size_t)x;
  return true;
}

// "a,b,c" -> {a, b, c} (src/local_cpu_multiway_merge_benchmark.cu:20-34).
inline bool ParseCountList(const std::string& s, std::vector<size_t>* out) {
  out->clear();
  size_t from = 0;
  while (true) {
    const size_t comma = s.find(',', from);
    size_t v = 0;
    if (!ParseCount(s.substr(from, comma == std::string::npos ? std::string::npos : comma - from), &v)) return false;
    out->push_back(v);
    if (comma == std::string::npos) return true;
    from = comma + 1;
  }
}

[[noreturn]] inline void UsageExit(const char* usage) {
  std::cout << usage;
  std::exit(1);
}

// The reference's failure line (printed after the CSV line), then exit 1.
inline void CheckOrExit(bool ok) {
  if (!ok) {
    std::cout << "Error: Invalid sort order.\n";
    std::exit(1);
  }
}

}  // namespace mgsort
