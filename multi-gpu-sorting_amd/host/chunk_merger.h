// The HET sort's host merge thread (MergeChunks of the reference,
// src/gowanlock_sort.cuh:20-95, with its SyncState counter :20-27): while the
// device still sorts, every group of `group_size` consecutive sorted chunks
// is merged into the secondary buffer as soon as the group's last chunk has
// landed in the primary buffer.  Host code only (no HIP), so the thread
// protocol can be checked on its own under ThreadSanitizer.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <mutex>
#include <parallel/algorithm>
#include <thread>
#include <utility>
#include <vector>

#include "key_traits.h"

namespace mgsort {

template <typename T>
void MultiwayMerge(std::vector<std::pair<T*, T*>>& runs, T* out, size_t len) {
  __gnu_parallel::multiway_merge(runs.begin(), runs.end(), out, (std::ptrdiff_t)len, EncodedLess<T>());
}

template <typename T>
class ChunkMerger {
 public:
  // Chunk c is primary[c chunk_size, c chunk_size + its length) (the last
  // one may be short: n bounds it); groups [0, groups) are merged, group g
  // covering chunks [g group_size, (g+1) group_size), into the same place of
  // `secondary`.  No thread is started when groups == 0.
  ChunkMerger(T* primary, T* secondary, size_t chunk_size, size_t n, size_t groups, size_t group_size)
      : primary_(primary), secondary_(secondary), cs_(chunk_size), n_(n), groups_(groups), mgs_(group_size) {
    if (groups_ > 0) thread_ = std::thread([this] { Loop(); });
  }
  ChunkMerger(const ChunkMerger&) = delete;
  ChunkMerger& operator=(const ChunkMerger&) = delete;
  // On an exception path: wakes the thread with every chunk marked landed
  // and joins it (its output is discarded with the error).
  ~ChunkMerger() {
    if (thread_.joinable()) {
      Landed(groups_ * mgs_);
      thread_.join();
    }
  }

  // Chunks [0, count) are sorted in the primary buffer.
  void Landed(size_t count) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      landed_ = std::max(landed_, count);
    }
    cv_.notify_one();
  }

  // Returns once every group is merged.
  void Join() {
    if (thread_.joinable()) thread_.join();
  }

 private:
  size_t ChunkLen(size_t c) const { return std::min(cs_, n_ - c * cs_); }

  void Loop() {
    for (size_t g = 0; g < groups_; ++g) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return landed_ >= (g + 1) * mgs_; });
      }
      std::vector<std::pair<T*, T*>> runs;
      size_t len = 0;
      const size_t first = g * mgs_;
      for (size_t c = first; c < first + mgs_; ++c) {
        runs.push_back({primary_ + c * cs_, primary_ + c * cs_ + ChunkLen(c)});
        len += ChunkLen(c);
      }
      MultiwayMerge(runs, secondary_ + first * cs_, len);
    }
  }

  T* primary_;
  T* secondary_;
  size_t cs_, n_, groups_, mgs_;
  std::mutex mu_;
  std::condition_variable cv_;
  size_t landed_ = 0;
  std::thread thread_;
};

}  // namespace mgsort
