// HET sort (Gowanlock & Karsin 2019) -- the reference's GowanlockSort<T>
// (src/gowanlock_sort.cuh:395-483) with its chunk and merge-group
// bookkeeping (:402-445, CalculateLastBufferParameters :97-115), on one
// MI355X.
//
// Pinned host chunks of chunk_size keys go H2D, are sorted in HBM and come
// back D2H; copies and sorts of consecutive chunks overlap on three HIP
// streams ordered by events.  num_buffers 2 (the reference's 2n plan):
// two chunk buffers, the sort of one using the other as scratch, H2D(c+1)
// overlapping D2H(c); num_buffers 3 (3n): two chunk buffers and the sort's
// scratch, the in-place transfer (D2H of chunk c-1 and H2D of chunk c+1
// through one buffer, block by block) overlapping the sort of chunk c.  A host thread
// merges groups of merge_group_size sorted runs with the GNU parallel
// multiway merge as they arrive (MergeChunks, :29-84), and a final multiway
// merge combines all runs (MergeMergeGroups, :86-95).
//
// Fix of reference defect D1: with merge_group_size == 1 every run goes to
// the final merge (the reference merges a zero-filled buffer instead).
//
// merge_backend "gpu": the sorted chunks stay resident in HBM (when the
// device has room for two copies of the input) and are combined by one k-way
// merge (mgs_multiway_merge), whose output goes back D2H in stretches while
// the later stretches merge; otherwise the CPU path runs.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "host_vector.h"

namespace mgsort {

struct GowanlockPlan {
  size_t chunk_size = 0;
  size_t merge_group_size = 0;
  size_t num_chunks = 0;
  size_t num_merge_groups_to_merge = 0;  // pre-merged by the host thread
  size_t num_chunks_to_merge = 0;
  size_t final_runs = 0;                 // runs of the final multiway merge
};

// Chunk and merge-group arithmetic of the reference for one GPU
// (src/gowanlock_sort.cuh:402-445), with the D1 fix.
GowanlockPlan PlanGowanlock(size_t num_elements, size_t chunk_size, size_t merge_group_size, size_t num_gpus);

void SetupGowanlockSort(const std::vector<int>& gpus, int num_threads);

// block_bytes: block of the in-place transfers of the 3-buffer schedule
// (0 = a hundredth of a chunk, as the reference's Core3n).
template <typename T>
void GowanlockSort(HostVector<T>* elements, size_t chunk_size, size_t merge_group_size, size_t num_buffers,
                   const std::vector<int>& gpus, const std::string& merge_backend = "cpu", size_t block_bytes = 0);

}  // namespace mgsort
