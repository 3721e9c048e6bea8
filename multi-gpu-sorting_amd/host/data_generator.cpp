// Seeded input generator (see data_generator.h).
#include "data_generator.h"

#include <omp.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <parallel/algorithm>
#include <random>
#include <type_traits>

namespace mgsort {

namespace {

constexpr size_t kNumBuckets = 10;

// SeedRandomGenerator of the reference (src/data_generator.cuh:193-205): a
// seed_seq of sizeof(mt19937::result_type) * 624 / sizeof(uint32) words, each
// the seed rotated left by one more bit than the previous.  On x86-64 glibc
// mt19937::result_type is 8 bytes wide, giving 1248 words.
std::mt19937 SeedMt(uint32_t s) {
  const size_t bytes = sizeof(std::mt19937::result_type) * std::mt19937::state_size;
  std::vector<std::seed_seq::result_type> words(bytes / sizeof(std::seed_seq::result_type));
  for (auto& w : words) {
    s = (s << 1) | (s >> 31);
    w = s;
  }
  std::seed_seq q(words.begin(), words.end());
  return std::mt19937{q};
}

// double -> T as the reference's static_cast behaves on x86-64: in-range
// values truncate; out-of-range values give INT_MIN / LONG_MIN (cvttsd2si)
// for the signed types, IEEE rounding (to +-inf) for float, and saturate for
// the added unsigned types.
template <typename T>
T FromDouble(double x) {
  if constexpr (std::is_same_v<T, float>) {
    return static_cast<float>(x);
  } else if constexpr (std::is_same_v<T, double>) {
    return x;
  } else if constexpr (std::is_signed_v<T>) {
    const double lim = std::ldexp(1.0, (int)sizeof(T) * 8 - 1);
    if (!(x >= -lim && x < lim)) return std::numeric_limits<T>::min();
    return static_cast<T>(x);
  } else {
    const double lim = std::ldexp(1.0, (int)sizeof(T) * 8);
    if (!(x > -1.0)) return 0;
    if (x >= lim) return std::numeric_limits<T>::max();
    return static_cast<T>(x);
  }
}

// a - b and comparisons of the nearly-sorted step in T's own arithmetic, with
// the two's-complement wrap x86 gives the reference's signed overflow.
template <typename T>
T WrapSub(T a, T b) {
  if constexpr (std::is_integral_v<T>) {
    using U = std::make_unsigned_t<T>;
    return static_cast<T>(static_cast<U>(static_cast<U>(a) - static_cast<U>(b)));
  } else {
    return a - b;
  }
}

template <typename T>
double MaxAsDouble() {
  return static_cast<double>(std::numeric_limits<T>::max());
}

// Runs fn(t, lo, hi) for each of `threads` logical generator threads over
// the static schedule of n iterations, on whatever cores OpenMP provides.
template <typename Fn>
void ForEachLogicalThread(size_t n, int threads, Fn fn) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < threads; ++t) {
    size_t lo, hi;
    DataGenerator::StaticRange(n, threads, t, &lo, &hi);
    fn(t, lo, hi);
  }
}

template <typename T>
void Uniform(T* b, size_t n, size_t seed, int threads) {
  ForEachLogicalThread(n, threads, [&](int t, size_t lo, size_t hi) {
    if (lo == hi) return;
    std::mt19937 g = SeedMt(static_cast<uint32_t>(seed + static_cast<size_t>(t)));
    std::uniform_real_distribution<double> d(0, MaxAsDouble<T>());
    for (size_t i = lo; i < hi; ++i) b[i] = FromDouble<T>(d(g));
  });
}

template <typename T>
void Normal(T* b, size_t n, size_t seed, int threads) {
  const double mean = MaxAsDouble<T>() / 2.0;
  const double stddev = mean / 3.0;
  ForEachLogicalThread(n, threads, [&](int t, size_t lo, size_t hi) {
    if (lo == hi) return;
    std::mt19937 g = SeedMt(static_cast<uint32_t>(seed + static_cast<size_t>(t)));
    std::normal_distribution<double> d(mean, stddev);
    for (size_t i = lo; i < hi; ++i) b[i] = FromDouble<T>(std::fabs(d(g)));
  });
}

template <typename T>
void Zero(T* b, size_t n) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) b[i] = 0;
}

// Bucket bounds [lower, upper] of the staggered distribution.
template <typename T>
std::uniform_real_distribution<double> StaggeredBucket(std::mt19937& g) {
  std::uniform_real_distribution<double> rr(0, MaxAsDouble<T>());
  const T upper = FromDouble<T>(rr(g));
  const T lower = (static_cast<double>(kNumBuckets) > static_cast<double>(upper))
                      ? T(0)
                      : WrapSub<T>(upper, static_cast<T>(kNumBuckets));
  return std::uniform_real_distribution<double>(static_cast<double>(lower), static_cast<double>(upper));
}

template <typename T>
void Staggered(T* b, size_t n, size_t seed, int threads) {
  if (n < kNumBuckets) {
    Uniform(b, n, seed, threads);
    return;
  }
  const size_t bs = n / kNumBuckets;
  const size_t iters = n / bs;
  ForEachLogicalThread(iters, threads, [&](int t, size_t lo, size_t hi) {
    if (lo == hi) return;
    std::mt19937 g = SeedMt(static_cast<uint32_t>(seed + static_cast<size_t>(t)));
    for (size_t i = lo; i < hi; ++i) {
      auto ud = StaggeredBucket<T>(g);
      for (size_t j = i * bs; j < (i + 1) * bs; ++j) b[j] = FromDouble<T>(ud(g));
    }
  });
  std::mt19937 g = SeedMt(static_cast<uint32_t>(seed + static_cast<size_t>(threads)));
  auto ud = StaggeredBucket<T>(g);
  for (size_t i = n - (n % bs); i < n; ++i) b[i] = FromDouble<T>(ud(g));
}

template <typename T>
void SortAscending(T* b, size_t n) {
  __gnu_parallel::sort(b, b + n);
}

template <typename T>
void Sorted(T* b, size_t n, size_t seed, int threads) {
  Uniform(b, n, seed, threads);
  SortAscending(b, n);
}

template <typename T>
void ReverseSorted(T* b, size_t n, size_t seed, int threads) {
  Uniform(b, n, seed, threads);
  __gnu_parallel::sort(b, b + n, std::greater<T>());
}

// Every gap is read from the sorted array (snapshot semantics).
template <typename T>
void NearlySorted(T* b, size_t n, size_t seed, int threads) {
  Sorted(b, n, seed, threads);
  if (n < 2) return;
  const size_t iters = n - 1;
  std::vector<T> boundary(threads);
  for (int t = 0; t < threads; ++t) {
    size_t lo, hi;
    DataGenerator::StaticRange(iters, threads, t, &lo, &hi);
    boundary[t] = b[hi];  // hi <= n - 1: still the sorted value
  }
  ForEachLogicalThread(iters, threads, [&](int t, size_t lo, size_t hi) {
    if (lo == hi) return;
    std::mt19937 g = SeedMt(static_cast<uint32_t>(seed + static_cast<size_t>(t)));
    for (size_t i = lo; i < hi; ++i) {
      const T next = (i + 1 < hi) ? b[i + 1] : boundary[t];
      double stddev = static_cast<double>(WrapSub<T>(next, b[i]));
      if (stddev < std::numeric_limits<double>::max() / 2.0) stddev *= 2.0;
      std::normal_distribution<double> d(0.0, stddev);
      const T diff = FromDouble<T>(std::fabs(d(g)));
      if (b[i] > WrapSub<T>(std::numeric_limits<T>::max(), diff))
        b[i] = std::numeric_limits<T>::max();
      else
        b[i] = static_cast<T>(b[i] + diff);
    }
  });
}

template <typename T>
void BucketSorted(T* b, size_t n, size_t seed, int threads) {
  const size_t bs = n / kNumBuckets;
  const size_t tail_begin = bs ? n - n % bs : 0;
  // Each logical thread draws its buckets (static schedule of 10 iterations),
  // then, from the same stream, its share of the tail.
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < threads; ++t) {
    size_t blo, bhi, tlo, thi;
    DataGenerator::StaticRange(kNumBuckets, threads, t, &blo, &bhi);
    DataGenerator::StaticRange(n - tail_begin, threads, t, &tlo, &thi);
    if (blo == bhi && tlo == thi) continue;
    std::mt19937 g = SeedMt(static_cast<uint32_t>(seed + static_cast<size_t>(t)));
    std::uniform_real_distribution<double> d(0, MaxAsDouble<T>());
    for (size_t i = blo; i < bhi; ++i)
      for (size_t j = i * bs; j < (i + 1) * bs; ++j) b[j] = FromDouble<T>(d(g));
    for (size_t i = tail_begin + tlo; i < tail_begin + thi; ++i) b[i] = FromDouble<T>(d(g));
  }
  // Elements between the tenth bucket and the tail (only when n / 10 <= n % 10)
  // are never drawn; the reference leaves its value-initialised zeros there.
  for (size_t i = kNumBuckets * bs; i < tail_begin; ++i) b[i] = 0;
  if (bs == 0) {
    SortAscending(b, n);
    return;
  }
  for (size_t i = 0; i <= kNumBuckets; ++i) {
    const size_t lo = i * bs, hi = std::min((i + 1) * bs, n);
    if (lo < hi) SortAscending(b + lo, hi - lo);
  }
}

// ---------------------------------------------------------------------------
// Added distributions: counter-based splitmix64 streams.
// ---------------------------------------------------------------------------
inline uint64_t SplitMix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
inline uint64_t StreamBase(uint32_t seed, uint32_t stream) {
  return SplitMix64((static_cast<uint64_t>(stream) << 32) | seed);
}
inline uint64_t Draw(uint64_t base, uint64_t i) { return SplitMix64(base + i); }
inline double Unit(uint64_t r) { return static_cast<double>(r >> 11) * 0x1.0p-53; }  // [0, 1)

template <typename T>
constexpr int Width() {
  return static_cast<int>(sizeof(T) * 8);
}

template <typename T>
void NormalClamped(T* b, size_t n, uint32_t seed) {
  const uint64_t base = StreamBase(seed, 1);
  double mean, sd, lo, hi;
  if constexpr (std::is_floating_point_v<T>) {
    mean = 0x1.0p31;
    sd = 0x1.0p28;
    lo = -std::numeric_limits<double>::infinity();
    hi = std::numeric_limits<double>::infinity();
  } else if constexpr (std::is_signed_v<T>) {
    mean = 0.0;
    sd = std::ldexp(1.0, Width<T>() - 4);
    lo = static_cast<double>(std::numeric_limits<T>::min());
    hi = static_cast<double>(std::numeric_limits<T>::max());
  } else {
    mean = std::ldexp(1.0, Width<T>() - 1);
    sd = std::ldexp(1.0, Width<T>() - 4);
    lo = 0.0;
    hi = static_cast<double>(std::numeric_limits<T>::max());
  }
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) {
    const double u1 = 1.0 - Unit(Draw(base, 2 * i));  // (0, 1]
    const double u2 = Unit(Draw(base, 2 * i + 1));
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    double x = mean + sd * z;
    if constexpr (std::is_floating_point_v<T>) {
      b[i] = static_cast<T>(x);
    } else {
      if (x <= lo) {
        b[i] = std::numeric_limits<T>::min();
      } else if (x >= hi) {
        b[i] = std::numeric_limits<T>::max();
      } else {
        b[i] = static_cast<T>(x);
      }
    }
  }
}

template <typename T>
T MapRank(uint64_t h) {
  if constexpr (std::is_floating_point_v<T>) {
    return static_cast<T>(static_cast<uint32_t>(h));
  } else {
    return static_cast<T>(h);
  }
}

template <typename T>
void Zipf(T* b, size_t n, uint32_t seed, double s) {
  constexpr size_t K = size_t(1) << 20;
  std::vector<double> cdf(K);
  double acc = 0.0;
  for (size_t k = 0; k < K; ++k) {
    acc += std::pow(static_cast<double>(k + 1), -s);
    cdf[k] = acc;
  }
  const double total = acc;
  const uint64_t base = StreamBase(seed, 2), vbase = StreamBase(seed, 3);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) {
    const double u = Unit(Draw(base, i)) * total;
    const size_t k = static_cast<size_t>(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    b[i] = MapRank<T>(Draw(vbase, std::min(k, K - 1)));
  }
}

template <typename T>
void NearlySortedSwaps(T* b, size_t n, uint32_t seed, int threads) {
  Sorted(b, n, seed, threads);
  if (n < 2) return;
  const uint64_t base = StreamBase(seed, 4);
  for (size_t j = 0; j < n / 100; ++j) {
    const size_t x = static_cast<size_t>(Draw(base, 2 * j) % n);
    const size_t y = static_cast<size_t>(Draw(base, 2 * j + 1) % n);
    std::swap(b[x], b[y]);
  }
}

template <typename T>
bool FloatSpecials(T* b, size_t n, uint32_t seed, int threads) {
  if constexpr (!std::is_floating_point_v<T>) {
    return false;
  } else {
    Uniform(b, n, seed, threads);
    using Bits = std::conditional_t<sizeof(T) == 4, uint32_t, uint64_t>;
    constexpr int kMant = sizeof(T) == 4 ? 23 : 52;
    constexpr Bits kSign = Bits(1) << (sizeof(T) * 8 - 1);
    constexpr Bits kExpAll = (sizeof(T) == 4) ? Bits(0x7F800000u) : Bits(0x7FF0000000000000ull);
    constexpr Bits kMantMask = (Bits(1) << kMant) - 1;
    const uint64_t base = StreamBase(seed, 5);
    const size_t count = std::min(n, std::max<size_t>(n / 100, 16));
    for (size_t j = 0; j < count; ++j) {
      const size_t pos = static_cast<size_t>(Draw(base, 2 * j) % n);
      const uint64_t r = Draw(base, 2 * j + 1);
      const Bits sign = (r & 1) ? kSign : Bits(0);
      const Bits payload = static_cast<Bits>(r >> 8) & kMantMask;
      Bits v;
      switch ((r >> 1) % 6) {
        case 0: v = sign; break;                                    // +-0
        case 1: v = sign | kExpAll; break;                          // +-inf
        case 2: v = sign | kExpAll | (Bits(1) << (kMant - 1)) | (payload >> 1); break;  // quiet NaN
        case 3: v = sign | kExpAll | (payload ? payload : Bits(1)); break;  // NaN (signalling or quiet)
        case 4: v = sign | (payload ? payload : Bits(1)); break;   // denormal
        default: v = sign | (Bits(1) << kMant) | payload; break;   // smallest-exponent normal
      }
      std::memcpy(&b[pos], &v, sizeof(T));
    }
    return true;
  }
}

}  // namespace

void DataGenerator::StaticRange(size_t n, int threads, int t, size_t* b, size_t* e) {
  const size_t nt = static_cast<size_t>(threads);
  size_t q = n / nt;
  size_t r = n % nt;
  const size_t tt = static_cast<size_t>(t);
  if (tt < r) {
    ++q;
    r = 0;
  }
  *b = q * tt + r;
  *e = *b + q;
}

const std::vector<std::string>& DataGenerator::ReferenceDistributions() {
  static const std::vector<std::string> v = {"uniform", "normal",         "zero",          "staggered",
                                             "sorted",  "reverse-sorted", "nearly-sorted", "bucket-sorted"};
  return v;
}

const std::vector<std::string>& DataGenerator::AddedDistributions() {
  static const std::vector<std::string> v = {"normal-clamped", "zipf", "zipf-1.2", "nearly-sorted-swaps",
                                             "float-specials"};
  return v;
}

bool DataGenerator::IsValidDistribution(const std::string& name) {
  const auto& a = ReferenceDistributions();
  const auto& b = AddedDistributions();
  return std::find(a.begin(), a.end(), name) != a.end() || std::find(b.begin(), b.end(), name) != b.end();
}

template <typename T>
bool DataGenerator::ComputeDistribution(T* begin, size_t n, const std::string& type) const {
  const size_t seed = distribution_seed_;
  const int th = num_threads_;
  const uint32_t s32 = static_cast<uint32_t>(seed);
  if (type == "uniform") Uniform(begin, n, seed, th);
  else if (type == "normal") Normal(begin, n, seed, th);
  else if (type == "zero") Zero(begin, n);
  else if (type == "staggered") Staggered(begin, n, seed, th);
  else if (type == "sorted") Sorted(begin, n, seed, th);
  else if (type == "reverse-sorted") ReverseSorted(begin, n, seed, th);
  else if (type == "nearly-sorted") NearlySorted(begin, n, seed, th);
  else if (type == "bucket-sorted") BucketSorted(begin, n, seed, th);
  else if (type == "normal-clamped") NormalClamped(begin, n, s32);
  else if (type == "zipf") Zipf(begin, n, s32, 1.0);
  else if (type == "zipf-1.2") Zipf(begin, n, s32, 1.2);
  else if (type == "nearly-sorted-swaps") NearlySortedSwaps(begin, n, s32, th);
  else if (type == "float-specials") return FloatSpecials(begin, n, s32, th);
  else return false;
  return true;
}

template bool DataGenerator::ComputeDistribution<int>(int*, size_t, const std::string&) const;
template bool DataGenerator::ComputeDistribution<long>(long*, size_t, const std::string&) const;
template bool DataGenerator::ComputeDistribution<float>(float*, size_t, const std::string&) const;
template bool DataGenerator::ComputeDistribution<double>(double*, size_t, const std::string&) const;
template bool DataGenerator::ComputeDistribution<uint32_t>(uint32_t*, size_t, const std::string&) const;
template bool DataGenerator::ComputeDistribution<uint64_t>(uint64_t*, size_t, const std::string&) const;

}  // namespace mgsort
