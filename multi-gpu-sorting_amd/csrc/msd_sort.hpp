// Three-pass hybrid radix sort of 4-byte keys (msd_sort.hip): shared
// geometry, control block and scratch layout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "internal.hpp"

namespace mgs {

constexpr int kMsdBins = 1 << 16;  // sub-buckets: the top 16 bits of the encoded key

struct MsdCfg {
  static constexpr int kBlock = 1024;  // one workgroup per CU
  static constexpr int kIpt = 18;      // pass tiles of 18432 keys
  static constexpr int kTile = kBlock * kIpt;
  // local sort: up to 20 keys per thread, 20480 keys (80 KB) in LDS (the
  // largest of the 65536 sub-buckets of 2^30 uniform keys holds about 17000)
  static constexpr int kLocalIpt = 20;
  static constexpr uint32_t kLocalCap = (uint32_t)kBlock * kLocalIpt;
  static constexpr uint32_t kMaxChains = 256;
  // sizes the path is tried for by default: [kMinKeys, kGapLo) and
  // [kGapHi, kMaxKeys], where it measured faster than the LSD sort on
  // uniform u32 and f32 keys (2^23-2^25: 2-9%; 2^29: 1%; 2^30: 21%); in
  // between, batches of several small sub-buckets take a third LDS pass and
  // the LSD sort is faster (2^26: 3%, 2^28: 12%).  MGS_MSD_MIN replaces both
  // ranges by [MGS_MSD_MIN, kMaxKeys].  kMaxKeys: the mean sub-bucket
  // (n / 65536) still leaves room below kLocalCap.
  static constexpr size_t kMinKeys = size_t(1) << 23;
  static constexpr size_t kGapLo = size_t(3) << 24;
  static constexpr size_t kGapHi = size_t(1) << 29;
  static constexpr size_t kMaxKeys = (size_t(1) << 30) + (size_t(1) << 27);
};

// Control block.  off: any set bit turns the path off (a 16-bit counter
// overflowed, a sub-bucket does not fit the local sort, a bucket is far
// above the mean); the LSD sort launched after the path runs iff off != 0.
// runs: the passes rank with the uniform-digit path (one LDS atomic for the
// lanes that share the first lane's digit), chosen when more than a
// sixteenth of the keys shared their top 16 bits with the first lane of
// their wavefront load in the upfront (runs of sorted keys, where most lanes
// of a wavefront would hit one counter).
// low: the sub-buckets are bits [low, low + 16) of the encoded key, where
// 16 - low is the number of leading bits (at most 16) that a sample of the
// keys shares; every key must share them, or the path is off.  Pass 1
// scatters by bits [low + 8, low + 16), pass 2 by bits [low, low + 8).
struct MsdCtl {
  uint32_t off;
  uint32_t batches;
  uint32_t batch_next;
  uint32_t xcd_mask;  // XCDs the upfront kernel ran on
  uint32_t runs;
  uint32_t low;
};

struct MsdLayout {
  size_t ctl, desc, slices, chain_hist, chain_pref, chain_bad, xcd_of, sub_cnt, sub_off, bucket_tot, bucket_start,
      leaders, batch_cnt, batch_base, batches, tile_base1, tile_base2, counters, end;
};

// n in the range the path is tried for (and MGS_MSD not set to 0).
bool msd_eligible(size_t n);
// The path's scratch from offset `at` of temp (it may overlap the LSD
// sort's descriptors, which only a sort the path turned itself off for
// uses), its control block at `ctl_at` (which must not).
MsdLayout msd_layout(size_t n, size_t at, size_t ctl_at);

// Enqueues the path's kernels: keys from kin, pass 1 into alt, pass 2 and
// the local sort into kout.  t is the temp buffer the layout indexes.
template <int KT>
int msd_launch(const void* kin, void* kout, void* alt, size_t n, char* t, const MsdLayout& L, uint32_t* status,
               uint32_t spin_limit, hipStream_t s, Marks& mark);

}  // namespace mgs
