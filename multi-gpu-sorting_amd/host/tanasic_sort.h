// P2P sort (Tanasic et al. 2013) -- the reference's TanasicSort<T>
// (src/tanasic_sort.cuh:223-299) and its helpers FindPivot (:45-87),
// SwapPartitions (:89-146), MergeLocalPartitions (:148-174) and the
// recursive MergePartitions (:176-204), on one MI355X.
//
// The reference gives each GPU one partition; here the partitions are
// `partitions` equal virtual slices of one device buffer (default: one per
// GPU, i.e. one), so the pivot search, the pairwise exchange and the merge of
// the two boundary partitions run exactly as on g GPUs, with the exchange
// done in place by mgs_swap_ranges instead of peer copies.  Several GPUs
// would need peer access between them; this build runs on one.
//
// Phases recorded in TimeDurations, as the reference: memory_allocation,
// sort_phase (H2D, fillers, per-partition sort, merge, D2H), merge_phase
// (nested in sort_phase), memory_deallocation.
#pragma once

#include <cstddef>
#include <vector>

#include "host_vector.h"

namespace mgsort {

void SetupTanasicSort(const std::vector<int>& gpus);

// Returns the pivot of the last (top-level) merge, 0 if there was none.
template <typename T>
size_t TanasicSort(HostVector<T>* elements, std::vector<int> gpus, size_t partitions = 0);

}  // namespace mgsort
