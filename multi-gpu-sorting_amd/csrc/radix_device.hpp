// Tile-level device routines of the radix sort, shared by the plain sort
// (radix_sort.hip) and the segmented sort (segmented.hip).
#pragma once

#include <type_traits>

#include "radix_sort.hpp"

namespace mgs {

#ifndef MGS_LB_SPIN_LIMIT
#define MGS_LB_SPIN_LIMIT (1u << 20)
#endif
constexpr uint32_t kLookbackSpinLimit = MGS_LB_SPIN_LIMIT;

// Predecessor descriptors read per look-back round trip: segmented sort
// tiles (2: fewer descriptor bytes per tile, measured 4-7% faster than 8),
// and the LSD pass kernel (whose window registers stay live across the digit
// scan and the scatter).
#ifndef MGS_LB_WINDOW
#define MGS_LB_WINDOW 2
#endif
#ifndef MGS_PASS_LB_WINDOW
#define MGS_PASS_LB_WINDOW 2
#endif

// Wavefront ranking used by every kernel: one LDS atomic per item by
// default; -DMGS_RANK_PEERS selects the peer-mask table instead.
#ifdef MGS_RANK_PEERS
#define MGS_RANK(d, peers, cnt) rank_wave((d), (peers), (cnt))
#else
#define MGS_RANK(d, peers, cnt) rank_wave_atomic((d), (cnt))
#endif

// The segmented kernels rank and count with the uniform-digit path (the
// lanes of a wavefront that share the first lane's digit take one atomic)
// unless MGS_SEG_UNI=0: measured on 4096 arrays of u64 keys + u32 payload,
// it costs 1.7% on uniform keys and saves 35% on keys with 1000 distinct
// values.
#ifndef MGS_SEG_UNI
#define MGS_SEG_UNI 1
#endif
constexpr bool kSegUni = MGS_SEG_UNI != 0;
// The segment histograms (every byte of every key counted once) count with
// MGS_SEG_HIST_UNI = 0: one atomic per key; 1: lds_count<true> (digit
// groups); 2: lds_count_wave (one atomic for a wavefront that shares one
// counter).
#ifndef MGS_SEG_HIST_UNI
#define MGS_SEG_HIST_UNI 2
#endif
__device__ __forceinline__ void seg_hist_count(uint32_t* c, uint32_t i) {
  if constexpr (MGS_SEG_HIST_UNI == 2)
    lds_count_wave(c, i);
  else
    lds_count<MGS_SEG_HIST_UNI != 0>(c, i);
}
#ifdef MGS_RANK_PEERS
#define MGS_SEG_RANK(d, peers, cnt) rank_wave((d), (peers), (cnt))
#else
#define MGS_SEG_RANK(d, peers, cnt) rank_wave_atomic<kSegUni>((d), (cnt))
#endif

constexpr int kWindow = MGS_LB_WINDOW;
constexpr int kPassWindow = MGS_PASS_LB_WINDOW;

// Issues the loads of the descriptors of predecessor tiles t, t-1, ...,
// t-NW+1 (those >= 0) for digit d of a chain whose first tile's
// descriptors start at `desc`.
template <typename D, int NW>
__device__ __forceinline__ void lookback_load(const D* desc, int64_t t, uint32_t d, D (&v)[NW]) {
#pragma unroll
  for (int j = 0; j < NW; ++j) v[j] = (t - j >= 0) ? ld_agent(desc + (size_t)(t - j) * kRadix + d) : D(0);
}

// Look-back for one digit starting from a window already loaded at
// predecessor t: sums aggregates, walking back, until an inclusive prefix is
// found (or the chain's first tile is passed).  Issuing the first window
// before the tile publishes its own aggregate keeps that store's completion
// off the critical path.
// Sums the leading published descriptors of a window loaded at predecessor
// t into `prefix`; returns how many were consumed and sets `done` when an
// inclusive prefix ended the walk.
template <typename D, int NW>
__device__ __forceinline__ int lookback_consume(int64_t t, int parity, const D (&v)[NW], uint32_t& prefix,
                                                bool& done) {
  int consumed = 0;
  done = false;
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    if (done || consumed != j || t - j < 0) continue;
    const uint32_t st = Desc<D>::status(v[j], parity);
    if (st == kStEmpty) continue;
    prefix += (uint32_t)Desc<D>::value(v[j]);
    consumed = j + 1;
    if (st == kStInclusive) done = true;
  }
  return consumed;
}

// One look-back round that does not wait for the next one: consumes the
// window loaded at predecessor t and, unless the walk is over, moves t back
// and issues the loads of the next window.  Returns true when the prefix is
// complete.
template <typename D, int NW>
__device__ __forceinline__ bool lookback_step(const D* desc, int64_t& t, uint32_t d, int parity, D (&v)[NW],
                                              uint32_t& prefix) {
  bool done;
  const int consumed = lookback_consume(t, parity, v, prefix, done);
  if (done || t - consumed < 0) return true;
  t -= consumed;
  lookback_load(desc, t, d, v);
  return false;
}

template <typename D, int NW>
__device__ __forceinline__ uint32_t lookback_from(const D* desc, int64_t t, uint32_t d, int parity,
                                                  D (&v)[NW], uint32_t prefix, uint32_t spin_limit, bool& gave_up) {
  uint32_t spins = 0;
  gave_up = false;
  for (;;) {
    bool done;
    const int consumed = lookback_consume(t, parity, v, prefix, done);
    if (done || t - consumed < 0) break;
    if (consumed == 0) {
      // Bounded wait: a predecessor tile always publishes within microseconds.
      // The bound keeps a corrupted descriptor array (or a chain no
      // workgroup serves) from hanging the GPU; the caller records it in the
      // status word, since the prefix is then incomplete (every store stays
      // in bounds).
      if (++spins > spin_limit) {
        gave_up = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    t -= consumed;
    lookback_load(desc, t, d, v);
  }
  return prefix;
}

// Look-back of tile `tile` (> 0) of a chain for digit d.
template <typename D>
__device__ __forceinline__ uint32_t lookback(const D* desc, uint32_t tile, uint32_t d, int parity,
                                             uint32_t spin_limit, bool& gave_up) {
  D v[kWindow];
  lookback_load(desc, (int64_t)tile - 1, d, v);
  return lookback_from(desc, (int64_t)tile - 1, d, parity, v, 0u, spin_limit, gave_up);
}

// ---------------------------------------------------------------------------
// One onesweep tile.
//
// A chain is a run of consecutive tiles whose digit counts are prefix-summed
// by look-back; the chain's starting offset of every digit is known before
// the pass (digit_base), so different chains never wait on each other.
//
// src/vsrc    the chain's input (index 0 = first key of the chain)
// n           keys in the chain
// dst/vdst    output arrays; digit_base indexes them directly
// n_dst       length of dst (bounds check of every store)
// tile        tile index inside the chain
// desc        the chain's descriptors (tile-major, 256 per tile)
// digit_base  256 output offsets of the chain's first key of each digit
// status      the call's status word (kStatusLookback when a wait gives up)
// ---------------------------------------------------------------------------
template <int KT, typename V, int BLOCK, int IPT, typename D>
__device__ __forceinline__ void onesweep_tile(const typename Codec<KT>::Bits* __restrict__ src,
                                              typename Codec<KT>::Bits* __restrict__ dst, const V* __restrict__ vsrc,
                                              V* __restrict__ vdst, uint32_t n, uint32_t n_dst, uint32_t tile,
                                              D* __restrict__ desc, const uint32_t* __restrict__ digit_base, int shift,
                                              int parity, uint32_t* __restrict__ status) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int W = BLOCK / kWave;
  constexpr int TILE = BLOCK * IPT;
  constexpr bool HV = kHasVal<V>;
  using VS = typename std::conditional<HV, V, uint32_t>::type;
  static_assert(W >= 4, "need >= 256 threads for the per-digit phase");

  struct RankSmem {
    uint64_t peers[W][kRadix];
  };
  union StageSmem {
    RankSmem r;
    Bits keys[TILE];
  };
  __shared__ StageSmem u;
  __shared__ VS vstage[HV ? TILE : 1];
  __shared__ uint32_t off[W][kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint32_t sums[4];

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;
  for (int i = tid; i < W * kRadix; i += BLOCK) {
    (&u.r.peers[0][0])[i] = 0;
    (&off[0][0])[i] = 0;
  }
  __syncthreads();
  const uint32_t tile_base = tile * (uint32_t)TILE;
  const uint32_t tile_cnt = min((uint32_t)TILE, n - tile_base);
  const bool full = tile_cnt == (uint32_t)TILE;

  Bits key[IPT];
  VS val[IPT];
  uint32_t dg[IPT], rk[IPT];
  const uint32_t wb = tile_base + (uint32_t)(w * IPT * kWave + lane);
  if (full) {
#pragma unroll
    for (int k = 0; k < IPT; ++k) key[k] = src[wb + k * kWave];
    if constexpr (HV) {
#pragma unroll
      for (int k = 0; k < IPT; ++k) val[k] = vsrc[wb + k * kWave];
    }
  } else {
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      const uint32_t idx = wb + k * kWave;
      key[k] = idx < n ? src[idx] : Bits(0);
      if constexpr (HV) val[k] = idx < n ? vsrc[idx] : VS(0);
    }
  }
#pragma unroll
  for (int k = 0; k < IPT; ++k) dg[k] = digit_of<KT>(key[k], shift);

  uint64_t* pt = u.r.peers[w];
  uint32_t* ct = off[w];
  (void)pt;
  if (full) {
#pragma unroll
    for (int k = 0; k < IPT; ++k) rk[k] = MGS_SEG_RANK(dg[k], pt, ct);
  } else {
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      if (wb + k * kWave < n) rk[k] = MGS_SEG_RANK(dg[k], pt, ct);
    }
  }
  __syncthreads();

  // Per-digit phase: threads 0..255 own one digit each.
  uint32_t total = 0;
  D* my_desc = desc + (size_t)tile * kRadix + tid;
  if (tid < kRadix) {
#pragma unroll
    for (int i = 0; i < W; ++i) total += off[i][tid];
    st_agent(my_desc, Desc<D>::make(tile == 0 ? kStInclusive : kStAggregate, (D)total, parity));
  }
  const uint32_t excl = block_excl_scan_256(total, sums, tid);
  if (tid < kRadix) {
    uint32_t run = excl;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      const uint32_t c = off[i][tid];
      off[i][tid] = run;
      run += c;
    }
    uint32_t prefix = 0;
    if (tile > 0) {
      bool gave_up;
      prefix = lookback(desc, tile, (uint32_t)tid, parity, kLookbackSpinLimit, gave_up);
      if (gave_up) atomicOr(status, kStatusLookback);
      st_agent(my_desc, Desc<D>::make(kStInclusive, (D)(prefix + total), parity));
    }
    gbase[tid] = digit_base[tid] + prefix - excl;
  }
  __syncthreads();

  // Tile-local scatter into LDS (sorted by digit, stable).
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    const uint32_t pos = off[w][dg[k]] + rk[k];
    if ((full || wb + k * kWave < n) && pos < (uint32_t)TILE) {
      u.keys[pos] = key[k];
      if constexpr (HV) vstage[pos] = val[k];
    }
  }
  __syncthreads();

  // Coalesced write-out: consecutive threads, consecutive positions.
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = (uint32_t)(j * BLOCK + tid);
    if (full || i < tile_cnt) {
      const Bits kk = u.keys[i];
      const uint32_t o = gbase[digit_of<KT>(kk, shift)] + i;
      if (o < n_dst) {
        dst[o] = kk;
        if constexpr (HV) vdst[o] = vstage[i];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Sorts n <= BLOCK*IPT keys (and payloads) with one workgroup in LDS.
// k_in may equal k_out.
// ---------------------------------------------------------------------------
template <int KT, typename V, int BLOCK, int IPT>
__device__ __forceinline__ void small_sort_block(const typename Codec<KT>::Bits* k_in, typename Codec<KT>::Bits* k_out,
                                                 const V* v_in, V* v_out, uint32_t n) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int W = BLOCK / kWave;
  constexpr int TILE = BLOCK * IPT;
  constexpr bool HV = kHasVal<V>;
  constexpr int P = (int)sizeof(Bits);
  using VS = typename std::conditional<HV, V, uint32_t>::type;

  __shared__ Bits keys[TILE];
  __shared__ VS vals[HV ? TILE : 1];
  __shared__ uint64_t peers[W][kRadix];
  __shared__ uint32_t off[W][kRadix];
  __shared__ uint32_t sums[4];
  __shared__ int s_flag;

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;
  for (int i = tid; i < W * kRadix; i += BLOCK) (&peers[0][0])[i] = 0;
  for (uint32_t i = tid; i < n; i += BLOCK) {
    keys[i] = k_in[i];
    if constexpr (HV) vals[i] = v_in[i];
  }
  __syncthreads();

  const uint32_t wb = (uint32_t)(w * IPT * kWave + lane);
  for (int p = 0; p < P; ++p) {
    const int shift = 8 * p;
    for (int i = tid; i < W * kRadix; i += BLOCK) (&off[0][0])[i] = 0;
    if (tid == 0) s_flag = 0;
    Bits key[IPT];
    VS val[IPT];
    uint32_t dg[IPT], rk[IPT];
    const uint32_t d_first = digit_of<KT>(keys[0], shift);
    bool mixed = false;
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      const uint32_t idx = wb + k * kWave;
      key[k] = idx < n ? keys[idx] : Bits(0);
      if constexpr (HV) val[k] = idx < n ? vals[idx] : VS(0);
      dg[k] = digit_of<KT>(key[k], shift);
      mixed |= (idx < n) && (dg[k] != d_first);
    }
    __syncthreads();
    if (mixed) s_flag = 1;
    __syncthreads();
    if (s_flag == 0) continue;  // every key has the same byte: identity pass
#pragma unroll
    for (int k = 0; k < IPT; ++k)
      if (wb + k * kWave < n) rk[k] = MGS_RANK(dg[k], peers[w], off[w]);
    __syncthreads();
    uint32_t total = 0;
    if (tid < kRadix) {
#pragma unroll
      for (int i = 0; i < W; ++i) total += off[i][tid];
    }
    const uint32_t excl = block_excl_scan_256(total, sums, tid);
    if (tid < kRadix) {
      uint32_t run = excl;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const uint32_t c = off[i][tid];
        off[i][tid] = run;
        run += c;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      const uint32_t pos = off[w][dg[k]] + rk[k];
      if (wb + k * kWave < n && pos < (uint32_t)TILE) {
        keys[pos] = key[k];
        if constexpr (HV) vals[pos] = val[k];
      }
    }
    __syncthreads();
  }
  for (uint32_t i = tid; i < n; i += BLOCK) {
    k_out[i] = keys[i];
    if constexpr (HV) v_out[i] = vals[i];
  }
}

}  // namespace mgs
