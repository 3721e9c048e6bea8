// PARADIS parallel in-place MSD radix sort (see paradis.h).
#include "paradis.h"

#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "key_traits.h"

namespace mgsort {

namespace {

constexpr int kBuckets = 256;
constexpr size_t kSmall = 64;          // insertion sort below this
constexpr size_t kSequential = 1 << 16;  // one thread below this

template <typename T>
inline uint32_t DigitOf(const T& v, int byte) {
  return static_cast<uint32_t>((Encode(v) >> (8 * byte)) & 0xFF);
}

template <typename T>
void InsertionSort(T* a, size_t n) {
  for (size_t i = 1; i < n; ++i) {
    T v = a[i];
    const auto ev = Encode(v);
    size_t j = i;
    for (; j > 0 && Encode(a[j - 1]) > ev; --j) a[j] = a[j - 1];
    a[j] = v;
  }
}

// Single-threaded in-place MSD radix sort (American flag sort) on bytes
// byte, byte-1, ..., 0.
template <typename T>
void FlagSort(T* a, size_t n, int byte) {
  if (n <= kSmall) {
    InsertionSort(a, n);
    return;
  }
  size_t cnt[kBuckets] = {0};
  for (size_t i = 0; i < n; ++i) cnt[DigitOf(a[i], byte)]++;
  size_t head[kBuckets], tail[kBuckets], s = 0;
  for (int b = 0; b < kBuckets; ++b) {
    head[b] = s;
    s += cnt[b];
    tail[b] = s;
  }
  for (int b = 0; b < kBuckets; ++b) {
    while (head[b] < tail[b]) {
      T v = a[head[b]];
      uint32_t k = DigitOf(v, byte);
      while (k != (uint32_t)b) {
        std::swap(v, a[head[k]++]);
        k = DigitOf(v, byte);
      }
      a[head[b]++] = v;
    }
  }
  if (byte == 0) return;
  size_t start = 0;
  for (int b = 0; b < kBuckets; ++b) {
    if (cnt[b] > 1) FlagSort(a + start, cnt[b], byte - 1);
    start += cnt[b];
  }
}

// One PARADIS level on a[0, n) for `byte` with p threads; afterwards every
// bucket holds exactly its own keys.  Returns the bucket sizes.
template <typename T>
void ParadisLevel(T* a, size_t n, int byte, int p, size_t* cnt) {
  // parallel histogram
  std::vector<size_t> local((size_t)p * kBuckets, 0);
#pragma omp parallel num_threads(p)
  {
    const int t = omp_get_thread_num();
    size_t* c = &local[(size_t)t * kBuckets];
    const size_t lo = n * t / p, hi = n * (t + 1) / p;
    for (size_t i = lo; i < hi; ++i) c[DigitOf(a[i], byte)]++;
  }
  for (int b = 0; b < kBuckets; ++b) {
    cnt[b] = 0;
    for (int t = 0; t < p; ++t) cnt[b] += local[(size_t)t * kBuckets + b];
  }
  // gh/gt: unfinished region [gh, gt) of each bucket
  std::vector<size_t> gh(kBuckets), gt(kBuckets);
  size_t s = 0;
  for (int b = 0; b < kBuckets; ++b) {
    gh[b] = s;
    s += cnt[b];
    gt[b] = s;
  }
  std::vector<size_t> ph((size_t)p * kBuckets), pt((size_t)p * kBuckets);
  size_t last_remaining = ~size_t(0);
  for (;;) {
    size_t remaining = 0;
    for (int b = 0; b < kBuckets; ++b) remaining += gt[b] - gh[b];
    if (remaining == 0) break;
    if (remaining >= last_remaining) p = 1;  // no progress: one thread finishes the level
    last_remaining = remaining;
    // stripes: thread t gets the t-th slice of every bucket's unfinished region
    for (int b = 0; b < kBuckets; ++b) {
      const size_t len = gt[b] - gh[b];
      for (int t = 0; t < p; ++t) {
        ph[(size_t)t * kBuckets + b] = gh[b] + len * t / p;
        pt[(size_t)t * kBuckets + b] = gh[b] + len * (t + 1) / p;
      }
    }
    // speculative permutation: each thread moves keys only inside its stripes
#pragma omp parallel num_threads(p)
    {
      const int t = omp_get_thread_num();
      size_t* h = &ph[(size_t)t * kBuckets];
      const size_t* e = &pt[(size_t)t * kBuckets];
      for (int i = 0; i < kBuckets; ++i) {
        size_t head = h[i];
        while (head < e[i]) {
          T v = a[head];
          uint32_t k = DigitOf(v, byte);
          while (k != (uint32_t)i && h[k] < e[k]) {
            std::swap(v, a[h[k]++]);
            k = DigitOf(v, byte);
          }
          if (k == (uint32_t)i) {
            a[head++] = a[h[i]];
            a[h[i]++] = v;
          } else {
            a[head++] = v;
          }
        }
      }
    }
    // repair: per bucket, swap the misplaced keys left in each thread's
    // stripe with own keys taken from the back, so [gh, tail) holds only the
    // bucket's keys and [tail, gt) only foreign ones (the next round's work)
#pragma omp parallel for num_threads(p) schedule(dynamic, 8)
    for (int i = 0; i < kBuckets; ++i) {
      size_t tail = gt[i];
      bool stop = false;
      for (int t = 0; t < p && !stop; ++t) {
        size_t head = ph[(size_t)t * kBuckets + i];
        while (head < std::min(pt[(size_t)t * kBuckets + i], tail)) {
          if (DigitOf(a[head], byte) == (uint32_t)i) {
            ++head;
            continue;
          }
          bool found = false;
          while (tail > head + 1) {
            --tail;
            if (DigitOf(a[tail], byte) == (uint32_t)i) {
              std::swap(a[head], a[tail]);
              found = true;
              break;
            }
          }
          if (!found) {  // everything from head on is foreign
            tail = head;
            stop = true;
            break;
          }
          ++head;
        }
      }
      gh[i] = tail;
    }
  }
}

template <typename T>
void ParadisRec(T* a, size_t n, int byte, int p) {
  if (n <= kSequential || p <= 1) {
    FlagSort(a, n, byte);
    return;
  }
  size_t cnt[kBuckets];
  ParadisLevel(a, n, byte, p, cnt);
  if (byte == 0) return;
  std::vector<size_t> start(kBuckets);
  size_t s = 0;
  for (int b = 0; b < kBuckets; ++b) {
    start[b] = s;
    s += cnt[b];
  }
  // large buckets: all threads, one after another; the rest: one thread each
  const size_t big = n / (size_t)p;
  for (int b = 0; b < kBuckets; ++b)
    if (cnt[b] > big) ParadisRec(a + start[b], cnt[b], byte - 1, p);
#pragma omp parallel for num_threads(p) schedule(dynamic, 1)
  for (int b = 0; b < kBuckets; ++b)
    if (cnt[b] > 1 && cnt[b] <= big) FlagSort(a + start[b], cnt[b], byte - 1);
}

}  // namespace

template <typename T>
void ParadisSort(T* begin, T* end, int num_threads) {
  const size_t n = static_cast<size_t>(end - begin);
  if (n < 2) return;
  ParadisRec(begin, n, (int)sizeof(T) - 1, std::max(1, num_threads));
}

template void ParadisSort<int32_t>(int32_t*, int32_t*, int);
template void ParadisSort<int64_t>(int64_t*, int64_t*, int);
template void ParadisSort<uint32_t>(uint32_t*, uint32_t*, int);
template void ParadisSort<uint64_t>(uint64_t*, uint64_t*, int);
template void ParadisSort<float>(float*, float*, int);
template void ParadisSort<double>(double*, double*, int);

}  // namespace mgsort
