// PARADIS: parallel in-place MSD radix sort (Cho, Brand, Bordawekar, Finkler,
// Kulandaisamy, Puri, "PARADIS: An Efficient Parallel Algorithm for In-place
// Radix Sort", VLDB 2015).  The reference's CPU baseline calls
// paradis::sort<T>(begin, end, threads) (src/sort_benchmark.cu:95); its
// third_party/simple_paradis source is not available, so this is a fresh
// implementation from the paper: per level, a parallel histogram, rounds of
// speculative permutation (each thread owns a stripe of every bucket) and
// repair until every bucket holds only its own keys, then recursion into the
// buckets on the next byte.  Keys are ordered by the library's encoded order
// (equal to operator< for NaN-free input).
#pragma once

#include <cstddef>

namespace mgsort {

template <typename T>
void ParadisSort(T* begin, T* end, int num_threads);

}  // namespace mgsort
