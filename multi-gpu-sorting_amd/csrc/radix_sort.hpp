// LSD onesweep radix sort for gfx950 (MI355X): shared definitions.
//
// Replaces the thrust::sort / cub::DeviceRadixSort call sites of the
// reference (src/tanasic_sort.cuh:264-266, src/gowanlock_sort.cuh:188-190,
// :337-340, src/sort_benchmark.cu:89, src/local_gpu_sort_algorithm_benchmark.cu
// :55-65).  Algorithm: Adinets & Merrill, "Onesweep" (2022) with the
// decoupled look-back of Merrill & Garland (2016), re-designed for 64-lane
// wavefronts and the eight L2 domains (XCDs) of MI355X: every pass splits
// its tiles into independent look-back chains, one per XCD, whose digit bases
// come from per-chain histograms counted by the previous pass.  The launch
// sequence is described in radix_sort.hip.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace mgs {

constexpr int kMaxPasses = 8;

// Status word: the first 32-bit word of every call's temp buffer, cleared by
// the call's first kernel.  Kernels set these bits when they detect a
// condition that leaves the output invalid; mgs_temp_status() reads it.
enum : uint32_t {
  kStatusLookback = 1u,        // a look-back wait reached its bound
  kStatusChainUnserved = 2u,   // a chain's tiles were not all taken
  kStatusSkipped = 4u,         // segments left unsorted (malformed bounds)
  kStatusUnsorted = 8u,        // multiway merge: runs not sorted (a piece over its bound)
};

// Buffer ids of the ping-pong plan.
enum : int32_t { kBufIn = 0, kBufAlt = 1, kBufOut = 2 };

// Look-back chains per pass (one per XCD by default).
#ifndef MGS_CHAINS
#define MGS_CHAINS 8
#endif
constexpr int kChains = MGS_CHAINS;

// Partition of a pass's tiles into chains: chain c holds the tiles
// [c*tpc, min((c+1)*tpc, tiles)), i.e. the keys [c*chain_keys, ...).
// chain_of() maps a key position to its chain with one multiply by a
// precomputed reciprocal (magic = ceil(2^40 / (chain_keys >> shift))) and one
// correction step, exact for every position below 2^32.
struct ChainMap {
  uint32_t tiles;
  uint32_t tpc;
  uint32_t chains;
  uint32_t shift;
  uint64_t chain_keys;
  uint64_t magic;
};

__host__ __device__ inline uint32_t chain_of(uint32_t pos, const ChainMap& m) {
  uint32_t c = (uint32_t)(((uint64_t)(pos >> m.shift) * m.magic) >> 40);
  if ((uint64_t)c * m.chain_keys > pos) --c;
  return c;
}

inline ChainMap make_chain_map(size_t n, uint32_t tile, int max_chains) {
  ChainMap m{};
  m.tiles = (uint32_t)((n + tile - 1) / tile);
  uint32_t ch = (uint32_t)std::max(1, std::min<int>(max_chains, (int)std::max<uint32_t>(m.tiles, 1u)));
  m.tpc = std::max<uint32_t>(1u, (m.tiles + ch - 1) / ch);
  m.chains = std::max<uint32_t>(1u, (m.tiles + m.tpc - 1) / m.tpc);
  m.shift = (uint32_t)__builtin_ctz(tile);
  m.chain_keys = (uint64_t)m.tpc * tile;
  const uint64_t d = m.chain_keys >> m.shift;
  m.magic = ((uint64_t(1) << 40) + d - 1) / d;
  return m;
}

// Look-back descriptor: 2-bit status code in the top bits, count below.
// The status code is rotated by the parity of the active pass index so that
// the final state of one pass (all "inclusive") reads as "empty" in the next,
// and the descriptor array is cleared only once per sort.
template <typename D> struct Desc {
  static constexpr int kShift = sizeof(D) * 8 - 2;
  static constexpr D kMask = (D(1) << kShift) - 1;
  __device__ static inline D make(uint32_t status, D value, int parity) {
    D code = (D)((status + 2u * (uint32_t)parity) & 3u);
    return (code << kShift) | (value & kMask);
  }
  __device__ static inline uint32_t status(D v, int parity) {
    uint32_t code = (uint32_t)(v >> kShift);
    return (code - 2u * (uint32_t)parity) & 3u;
  }
  __device__ static inline D value(D v) { return v & kMask; }
};
enum : uint32_t { kStEmpty = 0, kStAggregate = 1, kStInclusive = 2 };

// ---------------------------------------------------------------------------
// Tile geometry per (key bytes, value bytes).
// ---------------------------------------------------------------------------
#ifndef MGS_OS_BLOCK
#define MGS_OS_BLOCK 512
#endif
// Items per thread for 4-byte keys alone, for 8 bytes of key and payload,
// for 12, and for more.  18 is the most a u32 thread holds in 128 VGPRs
// (four wavefronts per SIMD) without spilling in the hot loop; the wider
// records run 512-thread workgroups, one per CU, with the largest tile the
// LDS holds beside the counters.
#ifndef MGS_OS_IPT
#define MGS_OS_IPT 18
#endif
// 4-byte keys alone run one 1024-thread workgroup per CU (16 wavefronts,
// 18432-key tiles): measured 5% faster at 2^30 keys than two 512-thread
// workgroups with 9216-key tiles, and 23% faster at 2^22.
#ifndef MGS_OS_BLOCK_4B
#define MGS_OS_BLOCK_4B 1024
#endif
#ifndef MGS_OS_IPT_8B
#define MGS_OS_IPT_8B 28
#endif
#ifndef MGS_OS_IPT_MID
#define MGS_OS_IPT_MID 16
#endif
#ifndef MGS_OS_IPT_WIDE
#define MGS_OS_IPT_WIDE 12
#endif
// Wavefronts per SIMD the LSD pass kernel is compiled for, which caps its
// VGPRs: four for 4-byte keys alone (one 1024-thread workgroup per CU, 128
// VGPRs), two for wider records, whose larger tiles need the registers and
// the LDS of a whole CU (measured at 2^28-2^29 keys: 16-21% faster for u32
// and u64 keys with a payload than two 512-thread workgroups with 8 or 4
// items per thread).  Arrays below kBigTileMin keys keep those small tiles
// and two workgroups per CU: they have too few tiles to fill the device
// otherwise.
#ifndef MGS_PASS_WAVES_4B
#define MGS_PASS_WAVES_4B 4
#endif
#ifndef MGS_PASS_WAVES_WIDE
#define MGS_PASS_WAVES_WIDE 2
#endif
template <int KB, int VB, int IPT4 = MGS_OS_IPT, int IPT8 = MGS_OS_IPT_8B, int IPT12 = MGS_OS_IPT_MID,
          int IPTW = MGS_OS_IPT_WIDE, int WAVES_WIDE = MGS_PASS_WAVES_WIDE, int BLOCK4 = MGS_OS_BLOCK_4B>
struct TileCfg {
  static constexpr int kBlock = (KB + VB) <= 4 ? BLOCK4 : MGS_OS_BLOCK;
  static constexpr int kIpt = (KB + VB) <= 4 ? IPT4 : (KB + VB) <= 8 ? IPT8 : (KB + VB) <= 12 ? IPT12 : IPTW;
  static constexpr int kTile = kBlock * kIpt;
  static constexpr int kWavesPerEu = (KB + VB) <= 4 ? MGS_PASS_WAVES_4B : WAVES_WIDE;
};

// Tile geometry of the LSD passes: large tiles, or the small ones below
// kBigTileMin keys (records wider than 4 bytes only; 4-byte keys have one
// geometry).
constexpr size_t kBigTileMin = size_t(1) << 21;
template <int KB, int VB, bool BIG>
using PassCfg = typename std::conditional<BIG || (KB + VB) <= 4, TileCfg<KB, VB>,
                                          TileCfg<KB, VB, MGS_OS_IPT, 8, 8, 4, MGS_PASS_WAVES_4B>>::type;
inline bool use_big_tiles(int kb, int vb, size_t n) { return kb + vb <= 4 || n >= kBigTileMin; }

// Single-workgroup LDS sort geometry.
template <int KB, int VB> struct SmallCfg {
  static constexpr int kBlock = 256;
  static constexpr int kIpt = (KB + VB) <= 8 ? 16 : 8;
  static constexpr int kTile = kBlock * kIpt;
};

// ---------------------------------------------------------------------------
// Stable per-wavefront ranking of one item per lane.
//
// peers:  this wavefront's 256-entry table of 64-bit lane masks (all zero on
//         entry and on exit)
// cnt:    this wavefront's running count per digit
// Returns the number of items with the same digit that precede this one in
// the wavefront's order (earlier items, then lower lanes of this item).
//
// Lanes with equal digits find each other through an LDS atomic OR of their
// lane bit into peers[d]; every lane then reads back the full mask, counts
// the lower lanes in it (mbcnt) and the lowest lane advances cnt[d].  DS
// instructions of one wavefront execute in order, so the read follows every
// lane's OR, and the clear follows every lane's read.  A wave whose active
// lanes all share one digit (runs in sorted input) skips the table.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t rank_wave(uint32_t d, uint64_t* peers, uint32_t* cnt) {
  const uint64_t act = ballot64(true);
  const uint32_t d0 = readfirstlane(d);
  const uint64_t same = ballot64(d == d0);
  if (same == act) {
    const uint32_t below = mbcnt64(act);
    const uint32_t base = ld_wg(&cnt[d0]);
    if (below == 0) st_wg(&cnt[d0], base + (uint32_t)__popcll(act));
    return base + below;
  }
  const int lane = lane_id();
  or_wg(&peers[d], (uint64_t)1 << lane);
  const uint64_t pm = ld_wg(&peers[d]);
  st_wg(&peers[d], (uint64_t)0);
  const uint32_t base = ld_wg(&cnt[d]);
  const uint32_t below = mbcnt64(pm);
  if (below == 0) st_wg(&cnt[d], base + (uint32_t)__popcll(pm));
  return base + below;
}

// Stable per-wavefront ranking with one LDS atomic per item: ds_add_rtn_u32
// on cnt[d] returns the count of earlier items with digit d.  The LDS serves
// the conflicting lanes of one instruction in ascending lane order (measured
// on gfx950 over 786k random conflict patterns; every multi-pass parity test
// with payloads re-checks it, since any other order breaks LSD stability).
// Tuning switches of the pass kernel (measured on MI355X; see the launch
// sequence in radix_sort.hip):
//   MGS_UNI        uniform-digit fast path in ranking and counting, in
//                  every pass of every sort
//   MGS_SKEW_PAIR_MIN  sorts of at least this many keys choose the ranking
//                  per pass from the digit counts (uniform-digit path when
//                  one digit holds more than n >> MGS_SKEW_SHIFT keys,
//                  1/64 of them by default)
//   MGS_XCD_LOCAL  chains served per XCD with workgroup-scope descriptor
//                  stores (when the device shows eight XCDs)
#ifndef MGS_XCD_LOCAL
#define MGS_XCD_LOCAL 1
#endif
#ifndef MGS_UNI
#define MGS_UNI 0
#endif
#ifndef MGS_SKEW_PAIR_MIN
#define MGS_SKEW_PAIR_MIN (1u << 26)
#endif
#ifndef MGS_SKEW_SHIFT
#define MGS_SKEW_SHIFT 6
#endif

// With UNI (the uniform-digit path), the lanes that carry the first active
// lane's digit (the frequent digit of heavy-duplicate keys, or the only one
// in runs of equal bytes) take one atomic for the whole group instead of one
// each: the first active lane adds the group's size, the other group lanes
// add 0 to a private slot of `sink`, and each group lane's rank is the first
// lane's result plus the number of group lanes below it.  The remaining lanes
// take one atomic each, on other counters.  Ranks stay in lane order within
// every digit.  Branch-free, so that the atomics of consecutive items issue
// back to back.
template <bool UNI = (MGS_UNI != 0)>
__device__ __forceinline__ uint32_t rank_wave_atomic(uint32_t d, uint32_t* cnt) {
  if constexpr (!UNI) {
  return __hip_atomic_fetch_add(&cnt[d], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
  __shared__ uint32_t sink[kWave];
  const uint32_t d0 = readfirstlane(d);
  const bool in = d == d0;  // the group of the first active lane's digit
  const uint64_t grp = ballot64(in);
  const uint32_t below = mbcnt64(grp);
  const bool lead = !in || below == 0;
  const uint32_t inc = in ? (below == 0 ? (uint32_t)__popcll(grp) : 0u) : 1u;
  const uint32_t r = __hip_atomic_fetch_add(lead ? &cnt[d] : &sink[lane_id()], inc, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_WORKGROUP);
  const uint32_t r0 = readfirstlane(r);
  return in ? r0 + below : r;
  }
}

// rank_wave_atomic for IPT items of a full wavefront at once: all atomics
// are issued first, the digit-group fix-ups follow, so the LDS round trips
// overlap.  Returns digit << 16 | rank per item.
template <int IPT, bool UNI = (MGS_UNI != 0)>
__device__ __forceinline__ void rank_items_full(const uint32_t (&d)[IPT], uint32_t* cnt, uint32_t (&pk)[IPT]) {
  if constexpr (!UNI) {
#pragma unroll
  for (int i = 0; i < IPT; ++i)
    pk[i] = (d[i] << 16) | __hip_atomic_fetch_add(&cnt[d[i]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
  __shared__ uint32_t sink[kWave];
  const uint32_t lane = lane_id();
  uint32_t r[IPT];
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const bool in = d[i] == readfirstlane(d[i]);  // lane 0's digit group
    const uint64_t grp = ballot64(in);
    const bool lead = !in || lane == 0;
    const uint32_t inc = in ? (lane == 0 ? (uint32_t)__popcll(grp) : 0u) : 1u;
    r[i] = __hip_atomic_fetch_add(lead ? &cnt[d[i]] : &sink[lane], inc, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_WORKGROUP);
  }
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const bool in = d[i] == readfirstlane(d[i]);
    const uint32_t rk = in ? readfirstlane(r[i]) + mbcnt64(ballot64(in)) : r[i];
    pk[i] = (d[i] << 16) | rk;
  }
  }
}

// Counts one item per active lane into LDS counter c[i]; with UNI, the lanes
// that count into the first active lane's counter add with one atomic for
// the group (same branch-free scheme as rank_wave_atomic).
template <bool UNI = (MGS_UNI != 0)>
__device__ __forceinline__ void lds_count(uint32_t* c, uint32_t i) {
  if constexpr (!UNI) {
  atomicAdd(&c[i], 1u);
  } else {
  __shared__ uint32_t sink[kWave];
  const bool in = i == readfirstlane(i);  // the group of the first active lane's counter
  const uint64_t grp = ballot64(in);
  const bool lead = !in || mbcnt64(grp) == 0;
  const uint32_t inc = in ? (lead ? (uint32_t)__popcll(grp) : 0u) : 1u;
  __hip_atomic_fetch_add(lead ? &c[i] : &sink[lane_id()], inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// Counts one item per active lane into LDS counter c[i]: one atomic when
// every active lane counts into the same counter (a wave-uniform branch),
// else one atomic per lane.  Cheaper per count than lds_count<true> when
// most wavefronts are mixed, as good when whole wavefronts share a counter.
__device__ __forceinline__ void lds_count_wave(uint32_t* c, uint32_t i) {
  const uint32_t i0 = readfirstlane(i);
  const uint64_t act = ballot64(true);
  if (ballot64(i == i0) == act) {
    if (mbcnt64(act) == 0) atomicAdd(&c[i0], (uint32_t)__popcll(act));
  } else {
    atomicAdd(&c[i], 1u);
  }
}

// Exclusive scan of one value per digit held by threads 0..255 (4 waves).
// `sums` is 4 words of LDS.  Must be called by every thread of the block
// (contains barriers); threads >= 256 pass 0 and get garbage.
__device__ __forceinline__ uint32_t block_excl_scan_256(uint32_t v, uint32_t* sums, int tid) {
  const uint32_t incl = wave_inclusive_sum(v);
  if (tid < 256 && (tid & 63) == 63) sums[tid >> 6] = incl;
  __syncthreads();
  uint32_t base = 0;
  const int w = tid >> 6;
#pragma unroll
  for (int i = 0; i < 4; ++i) base += (i < w) ? sums[i] : 0u;
  return base + incl - v;
}

}  // namespace mgs
