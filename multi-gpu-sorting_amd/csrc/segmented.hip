// Batched segmented sort: thousands of independent arrays in a fixed launch
// sequence whose grid sizes depend only on (nseg, max_total_keys), so a call
// can be captured once in a hipGraph and replayed for new data.
//
//   classify  one workgroup: resolves each segment's position (offsets form:
//             [offsets[s], offsets[s+1]); ranges form: [ranges[2s],
//             ranges[2s+1]) anywhere in the array, non-overlapping) and its
//             place in the compact scratch buffer, then bins it by length
//               small  (len <= small_max)  -> one workgroup sorts it in LDS
//               medium (len <= medium_max) -> one workgroup streams it through
//                                             HBM pass by pass, no look-back
//               large                      -> onesweep tiles chained by
//                                             decoupled look-back inside the
//                                             segment
//             and lays the large segments' tiles out back to back
//   clear     zeroes per-segment histograms and look-back descriptors
//   small     persistent grid over the small bin
//   medium    persistent grid over the medium bin
//   hist      per-tile histograms of the large segments
//   scan      per-segment digit bases
//   pass p    every large-segment tile of byte p in one launch
#include <type_traits>

#include "internal.hpp"
#include "radix_device.hpp"

namespace mgs {

namespace {

struct SegCtl {
  uint32_t status;  // the call's status word (first word of temp)
  uint32_t n_small, n_medium, n_large, large_tiles;
  uint32_t skipped;  // segments left unsorted: malformed bounds or beyond max_total
};

// Per-segment table written by classify: first key in the key array, first
// payload in the payload array, first slot in the compact scratch buffer,
// length.
struct SegTab {
  uint64_t* begin;
  uint64_t* vbegin;
  uint64_t* alt;
  uint32_t* len;
};

constexpr int kClassifyBlock = 1024;

// Exclusive scan of three counters across a 1024-thread block.
__device__ inline void block_scan3(uint32_t a, uint32_t b, uint32_t c, uint32_t& ea, uint32_t& eb, uint32_t& ec,
                                   uint32_t& ta, uint32_t& tb, uint32_t& tc, uint32_t (*sm)[16]) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t ia = wave_inclusive_sum(a), ib = wave_inclusive_sum(b), ic = wave_inclusive_sum(c);
  if (lane == 63) {
    sm[0][w] = ia;
    sm[1][w] = ib;
    sm[2][w] = ic;
  }
  __syncthreads();
  uint32_t ba = 0, bb = 0, bc = 0, sa = 0, sb = 0, sc = 0;
  for (int i = 0; i < kClassifyBlock / 64; ++i) {
    if (i < w) {
      ba += sm[0][i];
      bb += sm[1][i];
      bc += sm[2][i];
    }
    sa += sm[0][i];
    sb += sm[1][i];
    sc += sm[2][i];
  }
  ea = ba + ia - a;
  eb = bb + ib - b;
  ec = bc + ic - c;
  ta = sa;
  tb = sb;
  tc = sc;
  __syncthreads();
}

// Exclusive scan of 64-bit values across a 1024-thread block; *total gets
// the block sum.
__device__ inline uint64_t block_excl_scan_u64(uint64_t v, uint64_t* sm, uint64_t* total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint64_t incl = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint64_t t = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += t;
  }
  if (lane == 63) sm[w] = incl;
  __syncthreads();
  uint64_t base = 0, sum = 0;
  for (int i = 0; i < kClassifyBlock / 64; ++i) {
    if (i < w) base += sm[i];
    sum += sm[i];
  }
  *total = sum;
  __syncthreads();
  return base + incl - v;
}

// Inclusive running maximum of 64-bit values across a 1024-thread block;
// *total gets the block maximum.
__device__ inline uint64_t block_incl_max_u64(uint64_t v, uint64_t* sm, uint64_t* total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint64_t incl = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint64_t t = __shfl_up(incl, off, kWave);
    if (lane >= off) incl = incl > t ? incl : t;
  }
  if (lane == 63) sm[w] = incl;
  __syncthreads();
  uint64_t base = 0, all = 0;
  for (int i = 0; i < kClassifyBlock / 64; ++i) {
    if (i < w) base = base > sm[i] ? base : sm[i];
    all = all > sm[i] ? all : sm[i];
  }
  *total = all;
  __syncthreads();
  return incl > base ? incl : base;
}

template <bool RANGES>
__global__ __launch_bounds__(kClassifyBlock) void seg_classify_kernel(const uint64_t* __restrict__ bounds,
                                                                      const uint64_t* __restrict__ val_begin,
                                                                      uint32_t nseg, uint64_t max_total,
                                                                      uint64_t max_seg,
                                                                      uint32_t small_max, uint32_t medium_max,
                                                                      uint32_t tile, SegTab tab,
                                                                      uint32_t* small_list, uint32_t* medium_list,
                                                                      uint32_t* large_list, uint32_t* tile_start,
                                                                      SegCtl* ctl, uint32_t* tile_counter) {
  __shared__ uint32_t sm[3][16];
  __shared__ uint64_t sm64[16];
  const int tid = threadIdx.x;
  if (tid < kMaxPasses + 1) tile_counter[tid] = 0;
  uint32_t run_s = 0, run_m = 0, run_l = 0, run_t = 0, skipped = 0;
  uint64_t run_alt = 0, run_max = 0;
  const uint64_t base0 = RANGES ? 0 : bounds[0];
  for (uint32_t base = 0; base < nseg; base += kClassifyBlock) {
    const uint32_t s = base + tid;
    uint64_t b = 0, want = 0;
    bool ok = true;
    if (s < nseg) {
      const uint64_t lo = RANGES ? bounds[2 * (size_t)s] : bounds[s];
      const uint64_t hi = RANGES ? bounds[2 * (size_t)s + 1] : bounds[s + 1];
      b = lo;
      ok = hi >= lo && hi - lo < (uint64_t(1) << 32) && hi - lo <= max_seg;
      want = ok ? hi - lo : 0;
    }
    // compact scratch slot: the offsets form keeps the segment's distance
    // from the first offset; the ranges form packs the segments in order
    uint64_t alt = 0, chunk = 0;
    if (RANGES) {
      alt = run_alt + block_excl_scan_u64(want, sm64, &chunk);
      run_alt += chunk;
    } else {
      // offsets must not decrease: a segment is taken only if it starts at
      // or after every earlier offset, so the taken segments are disjoint
      // (b is at least every earlier offset iff the running maximum up to
      // and including it is b itself)
      const uint64_t m = block_incl_max_u64(s < nseg ? b : 0, sm64, &chunk);
      if (s < nseg && (m != b || b < run_max)) ok = false;
      run_max = run_max > chunk ? run_max : chunk;
      alt = b - base0;
    }
    if (s < nseg && (!ok || (!RANGES && b < base0) || alt > max_total || want > max_total - alt)) {
      ok = false;  // left unsorted rather than running past the scratch buffer
      want = 0;
    }
    const uint32_t len = (uint32_t)want;
    if (s < nseg) {
      tab.begin[s] = b;
      tab.vbegin[s] = (RANGES && val_begin != nullptr) ? val_begin[s] : b;
      tab.alt[s] = alt;
      tab.len[s] = len;
    }
    const uint32_t bad = (s < nseg && !ok) ? 1u : 0u;
    const uint32_t is_s = (s < nseg && len > 0 && len <= small_max) ? 1u : 0u;
    const uint32_t is_m = (s < nseg && len > small_max && len <= medium_max) ? 1u : 0u;
    const uint32_t is_l = (s < nseg && len > medium_max) ? 1u : 0u;
    const uint32_t nt = is_l ? (len + tile - 1) / tile : 0u;
    uint32_t es, em, el, ts, tm, tl;
    block_scan3(is_s, is_m, is_l, es, em, el, ts, tm, tl, sm);
    uint32_t et, eb, e2, tt, tb, t2;
    block_scan3(nt, bad, 0, et, eb, e2, tt, tb, t2, sm);
    if (is_s) small_list[run_s + es] = s;
    if (is_m) medium_list[run_m + em] = s;
    if (is_l) {
      large_list[run_l + el] = s;
      tile_start[run_l + el] = run_t + et;
    }
    run_s += ts;
    run_m += tm;
    run_l += tl;
    run_t += tt;
    skipped += tb;
  }
  if (tid == 0) {
    ctl->n_small = run_s;
    ctl->n_medium = run_m;
    ctl->n_large = run_l;
    ctl->large_tiles = run_t;
    ctl->skipped = skipped;
    ctl->status = skipped ? kStatusSkipped : 0u;
    tile_start[run_l] = run_t;
  }
}

__global__ void seg_clear_kernel(uint32_t* __restrict__ hist, uint32_t hist_words_per_seg,
                                 uint32_t* __restrict__ desc, uint32_t desc_words_per_tile,
                                 const SegCtl* __restrict__ ctl) {
  const size_t hw = (size_t)ctl->n_large * hist_words_per_seg;
  const size_t dw = (size_t)ctl->large_tiles * desc_words_per_tile;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  for (size_t i = g; i < hw; i += st) hist[i] = 0;
  for (size_t i = g; i < dw; i += st) desc[i] = 0;
}

template <int KT, typename V, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void seg_small_kernel(const typename Codec<KT>::Bits* k_in,
                                                          typename Codec<KT>::Bits* k_out, const V* v_in, V* v_out,
                                                          SegTab tab, const uint32_t* __restrict__ list,
                                                          const SegCtl* __restrict__ ctl) {
  const uint32_t cnt = ctl->n_small;
  for (uint32_t i = blockIdx.x; i < cnt; i += gridDim.x) {
    const uint32_t s = list[i];
    const uint64_t b = tab.begin[s];
    const uint32_t len = tab.len[s];
    const V* vi = nullptr;
    V* vo = nullptr;
    if constexpr (kHasVal<V>) {
      vi = v_in + tab.vbegin[s];
      vo = v_out + tab.vbegin[s];
    }
    small_sort_block<KT, V, BLOCK, IPT>(k_in + b, k_out + b, vi, vo, len);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Medium bin: one workgroup sorts a whole segment, pass by pass, walking its
// tiles in order and carrying the per-digit running offsets in LDS, so no
// cross-workgroup look-back is needed.
// ---------------------------------------------------------------------------
template <int KT, typename V, int BLOCK, int IPT>
__device__ void medium_sort_block(const typename Codec<KT>::Bits* k_in, typename Codec<KT>::Bits* k_alt,
                                  typename Codec<KT>::Bits* k_out, const V* v_in, V* v_alt, V* v_out, uint32_t n) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int W = BLOCK / kWave;
  constexpr int TILE = BLOCK * IPT;
  constexpr bool HV = kHasVal<V>;
  constexpr int P = (int)sizeof(Bits);
  using VS = typename std::conditional<HV, V, uint32_t>::type;
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
  __shared__ uint32_t hist[P][kRadix];
  __shared__ uint32_t run[kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint32_t sums[4];

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;

  // Histogram of every byte.
  for (int i = tid; i < P * kRadix; i += BLOCK) (&hist[0][0])[i] = 0;
  __syncthreads();
  for (uint32_t i = tid; i < n; i += BLOCK) {
    const Bits k = k_in[i];
#pragma unroll
    for (int p = 0; p < P; ++p) seg_hist_count(&hist[0][0], p * kRadix + digit_of<KT>(k, 8 * p));
  }
  __syncthreads();

  for (int p = 0; p < P; ++p) {
    const int shift = 8 * p;
    const Bits* src = p == 0 ? k_in : ((p & 1) ? k_alt : k_out);
    Bits* dst = (p & 1) ? k_out : k_alt;
    const V* vsrc = nullptr;
    V* vdst = nullptr;
    if constexpr (HV) {
      vsrc = p == 0 ? v_in : ((p & 1) ? v_alt : v_out);
      vdst = (p & 1) ? v_out : v_alt;
    }
    {  // running digit offsets start at the exclusive histogram scan
      const uint32_t c = tid < kRadix ? hist[p][tid] : 0u;
      const uint32_t e = block_excl_scan_256(c, sums, tid);
      if (tid < kRadix) run[tid] = e;
    }
    for (uint32_t tb = 0; tb < n; tb += TILE) {
      for (int i = tid; i < W * kRadix; i += BLOCK) {
        (&u.r.peers[0][0])[i] = 0;
        (&off[0][0])[i] = 0;
      }
      __syncthreads();
      const uint32_t tile_cnt = min((uint32_t)TILE, n - tb);
      Bits key[IPT];
      VS val[IPT];
      uint32_t dg[IPT], rk[IPT];
      const uint32_t wb = tb + (uint32_t)(w * IPT * kWave + lane);
#pragma unroll
      for (int k = 0; k < IPT; ++k) {
        const uint32_t idx = wb + k * kWave;
        key[k] = idx < n ? src[idx] : Bits(0);
        if constexpr (HV) val[k] = idx < n ? vsrc[idx] : VS(0);
        dg[k] = digit_of<KT>(key[k], shift);
      }
#pragma unroll
      for (int k = 0; k < IPT; ++k)
        if (wb + k * kWave < n) rk[k] = MGS_SEG_RANK(dg[k], u.r.peers[w], off[w]);
      __syncthreads();
      uint32_t total = 0;
      if (tid < kRadix) {
#pragma unroll
        for (int i = 0; i < W; ++i) total += off[i][tid];
      }
      const uint32_t excl = block_excl_scan_256(total, sums, tid);
      if (tid < kRadix) {
        uint32_t r = excl;
#pragma unroll
        for (int i = 0; i < W; ++i) {
          const uint32_t c = off[i][tid];
          off[i][tid] = r;
          r += c;
        }
        gbase[tid] = run[tid] - excl;
        run[tid] += total;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < IPT; ++k) {
        const uint32_t pos = off[w][dg[k]] + rk[k];
        if (wb + k * kWave < n && pos < (uint32_t)TILE) {
          u.keys[pos] = key[k];
          if constexpr (HV) vstage[pos] = val[k];
        }
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t i = (uint32_t)(j * BLOCK + tid);
        if (i < tile_cnt) {
          const Bits kk = u.keys[i];
          const uint32_t o = gbase[digit_of<KT>(kk, shift)] + i;
          if (o < n) {
            dst[o] = kk;
            if constexpr (HV) vdst[o] = vstage[i];
          }
        }
      }
      __syncthreads();
    }
    __threadfence_block();
    __syncthreads();
  }
}

template <int KT, typename V, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void seg_medium_kernel(const typename Codec<KT>::Bits* k_in,
                                                           typename Codec<KT>::Bits* k_alt,
                                                           typename Codec<KT>::Bits* k_out, const V* v_in, V* v_alt,
                                                           V* v_out, SegTab tab, const uint32_t* __restrict__ list,
                                                           const SegCtl* __restrict__ ctl) {
  const uint32_t cnt = ctl->n_medium;
  for (uint32_t i = blockIdx.x; i < cnt; i += gridDim.x) {
    const uint32_t s = list[i];
    const uint64_t b = tab.begin[s];
    const uint32_t len = tab.len[s];
    const uint64_t ab = tab.alt[s];
    const V* vi = nullptr;
    V *va = nullptr, *vo = nullptr;
    if constexpr (kHasVal<V>) {
      const uint64_t vb = tab.vbegin[s];
      vi = v_in + vb;
      va = v_alt + ab;
      vo = v_out + vb;
    }
    medium_sort_block<KT, V, BLOCK, IPT>(k_in + b, k_alt + ab, k_out + b, vi, va, vo, len);
    __syncthreads();
  }
}

// Large bin: locate tile g among the large segments.
__device__ __forceinline__ uint32_t find_large(const uint32_t* tile_start, uint32_t n_large, uint32_t g) {
  uint32_t lo = 0, hi = n_large;  // largest li with tile_start[li] <= g
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (tile_start[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <int KT, int BLOCK, int TILE>
__global__ __launch_bounds__(BLOCK) void seg_hist_kernel(const typename Codec<KT>::Bits* __restrict__ k_in,
                                                         SegTab tab, const uint32_t* __restrict__ large_list,
                                                         const uint32_t* __restrict__ tile_start,
                                                         uint32_t* __restrict__ seg_hist,
                                                         const SegCtl* __restrict__ ctl) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int P = (int)sizeof(Bits);
  __shared__ uint32_t h[P][kRadix];
  const uint32_t g = blockIdx.x;
  if (g >= ctl->large_tiles) return;
  const int tid = threadIdx.x;
  for (int i = tid; i < P * kRadix; i += BLOCK) (&h[0][0])[i] = 0;
  __syncthreads();
  const uint32_t li = find_large(tile_start, ctl->n_large, g);
  const uint32_t s = large_list[li];
  const uint64_t b = tab.begin[s];
  const uint32_t len = tab.len[s];
  const uint32_t t0 = (g - tile_start[li]) * (uint32_t)TILE;
  const uint32_t t1 = min(len, t0 + (uint32_t)TILE);
  for (uint32_t i = t0 + tid; i < t1; i += BLOCK) {
    const Bits k = k_in[b + i];
#pragma unroll
    for (int p = 0; p < P; ++p) seg_hist_count(&h[0][0], p * kRadix + digit_of<KT>(k, 8 * p));
  }
  __syncthreads();
  uint32_t* dsth = seg_hist + (size_t)li * P * kRadix;
  for (int i = tid; i < P * kRadix; i += BLOCK) {
    const uint32_t c = (&h[0][0])[i];
    if (c) atomicAdd(&dsth[i], c);
  }
}

template <int P>
__global__ __launch_bounds__(256) void seg_scan_kernel(const uint32_t* seg_hist, uint32_t* seg_base,
                                                       const SegCtl* __restrict__ ctl) {
  __shared__ uint32_t sums[4];
  const uint32_t li = blockIdx.x;
  if (li >= ctl->n_large) return;
  const int tid = threadIdx.x;
  for (int p = 0; p < P; ++p) {
    const size_t o = ((size_t)li * P + p) * kRadix + tid;
    const uint32_t c = seg_hist[o];
    const uint32_t e = block_excl_scan_256(c, sums, tid);
    __syncthreads();
    seg_base[o] = e;  // may alias seg_hist: every read of this row is done
  }
}

template <int KT, typename V, int BLOCK, int IPT, typename D>
__global__ __launch_bounds__(BLOCK) void seg_pass_kernel(const typename Codec<KT>::Bits* k_in,
                                                         typename Codec<KT>::Bits* k_alt,
                                                         typename Codec<KT>::Bits* k_out, const V* v_in, V* v_alt,
                                                         V* v_out, SegTab tab,
                                                         const uint32_t* __restrict__ large_list,
                                                         const uint32_t* __restrict__ tile_start,
                                                         const uint32_t* __restrict__ seg_base,
                                                         uint32_t* __restrict__ tile_counter, D* __restrict__ desc,
                                                         const SegCtl* __restrict__ ctl, uint32_t* __restrict__ status,
                                                         int pass) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int P = (int)sizeof(Bits);
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(&tile_counter[pass], 1u);
  __syncthreads();
  const uint32_t g = s_tile;
  if (g >= ctl->large_tiles) return;
  const uint32_t li = find_large(tile_start, ctl->n_large, g);
  const uint32_t s = large_list[li];
  const uint64_t b = tab.begin[s];
  const uint64_t ab = tab.alt[s];
  const uint32_t len = tab.len[s];
  const Bits* src;
  Bits* dst;
  if (pass == 0) {
    src = k_in + b;
    dst = k_alt + ab;
  } else if (pass & 1) {
    src = k_alt + ab;
    dst = k_out + b;
  } else {
    src = k_out + b;
    dst = k_alt + ab;
  }
  const V* vsrc = nullptr;
  V* vdst = nullptr;
  if constexpr (kHasVal<V>) {
    const uint64_t vb = tab.vbegin[s];
    if (pass == 0) {
      vsrc = v_in + vb;
      vdst = v_alt + ab;
    } else if (pass & 1) {
      vsrc = v_alt + ab;
      vdst = v_out + vb;
    } else {
      vsrc = v_out + vb;
      vdst = v_alt + ab;
    }
  }
  const uint32_t t0 = tile_start[li];
  onesweep_tile<KT, V, BLOCK, IPT, D>(src, dst, vsrc, vdst, len, len, g - t0, desc + (size_t)t0 * kRadix,
                                      seg_base + ((size_t)li * P + pass) * kRadix, pass * 8, pass & 1, status);
}

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct SegLayout {
  size_t alt_k, alt_v, ctl, tab_begin, tab_vbegin, tab_alt, tab_len, small_list, medium_list, large_list, tile_start, seg_hist,
      seg_base, tile_counter, desc, total;
  uint32_t max_large, tiles_ub;
  bool wide_desc;
};

// Items per thread of the large segments' tiles for 8- and 12-byte records
// and for 16-byte ones (one workgroup per CU; 20 and 14 measured 24% faster
// than 8 and 4 on 4096 arrays of u64 keys with a u32 payload).
#ifndef MGS_SEG_IPT_WIDE
#define MGS_SEG_IPT_WIDE 20
#endif
#ifndef MGS_SEG_IPT_WIDE16
#define MGS_SEG_IPT_WIDE16 14
#endif

// Tile geometries of the large segments' passes, chosen per call by
// mgs_tuning::large_variant (0 = the default; the autotuner times all of
// them per record width, payload and size):
//   record   variant 0        variant 1          variant 2
//   4 B      512 x 16         1024 x 18          512 x 8
//   8 B      512 x 20         512 x 28           512 x 12
//   12 B     512 x 20         512 x 16           512 x 12
//   16 B     512 x 14         512 x 12           512 x 8
// (threads x items per thread).  Variant 0: 16 items per thread for 4-byte
// keys (two workgroups per CU); wider records use the LDS of a whole CU, as
// the plain sort's large tiles do.  Variant 1 of the 4-byte keys is the
// plain sort's tile; the others trade tile size for occupancy.
constexpr int kSegVariants = 3;
template <int KB, int VB, int VAR>
struct SegLargeGeom {
  static constexpr int kRec = KB + VB;
  static constexpr int kBlock = (kRec <= 4 && VAR == 1) ? 1024 : 512;
  static constexpr int kIpt = kRec <= 4    ? (VAR == 0 ? 16 : VAR == 1 ? 18 : 8)
                              : kRec <= 8  ? (VAR == 0 ? MGS_SEG_IPT_WIDE : VAR == 1 ? 28 : 12)
                              : kRec <= 12 ? (VAR == 0 ? MGS_SEG_IPT_WIDE : VAR == 1 ? 16 : 12)
                                           : (VAR == 0 ? MGS_SEG_IPT_WIDE16 : VAR == 1 ? 12 : 8);
  static constexpr int kTile = kBlock * kIpt;
};

template <int KB, int VB>
inline int seg_large_tile(int var) {
  switch (var) {
    case 1: return SegLargeGeom<KB, VB, 1>::kTile;
    case 2: return SegLargeGeom<KB, VB, 2>::kTile;
    default: return SegLargeGeom<KB, VB, 0>::kTile;
  }
}

template <int KB, int VB>
struct SegCfg {
  // one workgroup per medium segment streams it through tiles of this size
  using M = TileCfg<KB, VB, 16, 8, 8, 4, 4, MGS_OS_BLOCK>;
  using S = SmallCfg<KB, VB>;
  static constexpr int kSmallMax = S::kTile;
  static constexpr int kMediumMax = 1 << 17;
};

// Large segments always hold more than kMinLargeExclusive keys, whatever the
// tuning, so the scratch size depends only on (max_total, nseg).
constexpr uint32_t kMinLargeExclusive = 256;

template <int KB, int VB>
SegLayout seg_layout(size_t max_total, size_t nseg, size_t max_seg, int large_tile) {
  constexpr int P = KB;
  SegLayout L{};
  L.max_large = (uint32_t)std::min<size_t>(nseg, max_total / (kMinLargeExclusive + 1) + 1);
  L.tiles_ub = (uint32_t)(max_total / (size_t)large_tile + L.max_large + 1);
  // descriptor values are prefixes inside one segment: 30 bits hold them
  // unless a segment may reach 2^30 keys
  L.wide_desc = max_seg >= (size_t(1) << 30);
  size_t o = 0;
  L.ctl = o;  // first: its status word is the call's status word
  o = align_up(o + sizeof(SegCtl));
  L.alt_k = o;
  o = align_up(o + max_total * KB);
  L.alt_v = o;
  o = align_up(o + max_total * VB);
  L.tab_begin = o;
  o = align_up(o + nseg * 8);
  L.tab_vbegin = o;
  o = align_up(o + nseg * 8);
  L.tab_alt = o;
  o = align_up(o + nseg * 8);
  L.tab_len = o;
  o = align_up(o + nseg * 4);
  L.small_list = o;
  o = align_up(o + nseg * 4);
  L.medium_list = o;
  o = align_up(o + nseg * 4);
  L.large_list = o;
  o = align_up(o + nseg * 4);
  L.tile_start = o;
  o = align_up(o + (nseg + 1) * 4);
  L.seg_hist = o;  // histograms, scanned in place into digit bases
  o = align_up(o + (size_t)L.max_large * P * kRadix * 4);
  L.seg_base = L.seg_hist;
  L.tile_counter = o;
  o = align_up(o + (kMaxPasses + 1) * 4);
  L.desc = o;
  o = align_up(o + (size_t)L.tiles_ub * kRadix * (L.wide_desc ? 8 : 4));
  L.total = o;
  return L;
}

// Longest segment the call accepts: the caller's hint (mgs_tuning::
// max_segment_keys) or max_total.  Longer segments are left unsorted and
// reported in the status word.
inline size_t segment_bound(size_t max_total, const mgs_tuning* tuning) {
  if (tuning != nullptr && tuning->max_segment_keys > 0)
    return std::min<size_t>(max_total, (size_t)tuning->max_segment_keys);
  return max_total;
}

inline int large_variant_of(const mgs_tuning* tuning) { return tuning != nullptr ? tuning->large_variant : 0; }

// The large segments' histogram, scan and passes with the tile geometry of
// variant VAR.
template <int KT, int VT, int VAR>
void seg_large(const typename Codec<KT>::Bits* ki, typename Codec<KT>::Bits* ka, typename Codec<KT>::Bits* ko,
               const typename ValOf<VT>::T* vi, typename ValOf<VT>::T* va, typename ValOf<VT>::T* vo,
               const SegTab& tab, const uint32_t* large_list, const uint32_t* tile_start, uint32_t* seg_hist,
               uint32_t* seg_base, uint32_t* tcount, void* desc, SegCtl* ctl, const SegLayout& L, hipStream_t s) {
  using V = typename ValOf<VT>::T;
  constexpr int KB = sizeof(typename Codec<KT>::Bits);
  constexpr int VB = kValBytes<V>;
  constexpr int P = KB;
  using G = SegLargeGeom<KB, VB, VAR>;
  seg_hist_kernel<KT, 256, G::kTile><<<L.tiles_ub, 256, 0, s>>>(ki, tab, large_list, tile_start, seg_hist, ctl);
  seg_scan_kernel<P><<<L.max_large, 256, 0, s>>>(seg_hist, seg_base, ctl);
  for (int p = 0; p < P; ++p) {
    if (L.wide_desc)
      seg_pass_kernel<KT, V, G::kBlock, G::kIpt, uint64_t><<<L.tiles_ub, G::kBlock, 0, s>>>(
          ki, ka, ko, vi, va, vo, tab, large_list, tile_start, seg_base, tcount, (uint64_t*)desc, ctl, &ctl->status, p);
    else
      seg_pass_kernel<KT, V, G::kBlock, G::kIpt, uint32_t><<<L.tiles_ub, G::kBlock, 0, s>>>(
          ki, ka, ko, vi, va, vo, tab, large_list, tile_start, seg_base, tcount, (uint32_t*)desc, ctl, &ctl->status, p);
  }
}

template <int KT, int VT>
int seg_typed(const void* kin, void* kout, const void* vin, void* vout, const uint64_t* bounds, bool ranges,
              const uint64_t* val_begin, size_t nseg, size_t max_total, const mgs_tuning* tuning, void* temp, size_t temp_bytes, hipStream_t s) {
  using Bits = typename Codec<KT>::Bits;
  using V = typename ValOf<VT>::T;
  constexpr int KB = sizeof(Bits);
  constexpr int VB = kValBytes<V>;
  constexpr int P = KB;
  using C = SegCfg<KB, VB>;
  using SC = typename C::S;
  using VP = typename std::conditional<kHasVal<V>, V, uint32_t>::type;

  uint32_t small_max = C::kSmallMax, medium_max = C::kMediumMax;
  if (tuning) {
    if (tuning->small_max > 0) small_max = std::min<uint32_t>((uint32_t)tuning->small_max, C::kSmallMax);
    if (tuning->medium_max > 0) medium_max = (uint32_t)tuning->medium_max;
  }
  medium_max = std::max({medium_max, small_max, kMinLargeExclusive});
  const size_t max_seg = segment_bound(max_total, tuning);
  const int var = large_variant_of(tuning);
  if (var < 0 || var >= kSegVariants) return MGS_ERR_INVALID;
  const int large_tile = seg_large_tile<KB, VB>(var);
  const SegLayout L = seg_layout<KB, VB>(max_total, nseg, max_seg, large_tile);
  if (temp == nullptr || temp_bytes < L.total) return MGS_ERR_TEMP_TOO_SMALL;
  char* t = static_cast<char*>(temp);
  Bits* ka = reinterpret_cast<Bits*>(t + L.alt_k);
  VP* va = reinterpret_cast<VP*>(t + L.alt_v);
  SegCtl* ctl = reinterpret_cast<SegCtl*>(t + L.ctl);
  uint32_t* small_list = reinterpret_cast<uint32_t*>(t + L.small_list);
  uint32_t* medium_list = reinterpret_cast<uint32_t*>(t + L.medium_list);
  uint32_t* large_list = reinterpret_cast<uint32_t*>(t + L.large_list);
  uint32_t* tile_start = reinterpret_cast<uint32_t*>(t + L.tile_start);
  uint32_t* seg_hist = reinterpret_cast<uint32_t*>(t + L.seg_hist);
  uint32_t* seg_base = reinterpret_cast<uint32_t*>(t + L.seg_base);
  uint32_t* tcount = reinterpret_cast<uint32_t*>(t + L.tile_counter);
  void* desc = t + L.desc;
  const SegTab tab{reinterpret_cast<uint64_t*>(t + L.tab_begin), reinterpret_cast<uint64_t*>(t + L.tab_vbegin),
                   reinterpret_cast<uint64_t*>(t + L.tab_alt), reinterpret_cast<uint32_t*>(t + L.tab_len)};

  const Bits* ki = static_cast<const Bits*>(kin);
  Bits* ko = static_cast<Bits*>(kout);
  const V* vi = static_cast<const V*>(vin);
  V* vo = static_cast<V*>(vout);

  if (ranges)
    seg_classify_kernel<true><<<1, kClassifyBlock, 0, s>>>(bounds, val_begin, (uint32_t)nseg, max_total, max_seg, small_max, medium_max,
                                                           (uint32_t)large_tile, tab, small_list, medium_list, large_list,
                                                           tile_start, ctl, tcount);
  else
    seg_classify_kernel<false><<<1, kClassifyBlock, 0, s>>>(bounds, val_begin, (uint32_t)nseg, max_total, max_seg, small_max, medium_max,
                                                            (uint32_t)large_tile, tab, small_list, medium_list, large_list,
                                                            tile_start, ctl, tcount);
  seg_clear_kernel<<<2048, 256, 0, s>>>(seg_hist, P * kRadix, (uint32_t*)desc, kRadix * (L.wide_desc ? 2 : 1), ctl);
  {
    const unsigned grid = (unsigned)std::min<size_t>(nseg, 8192);
    seg_small_kernel<KT, V, SC::kBlock, SC::kIpt><<<grid, SC::kBlock, 0, s>>>(ki, ko, vi, vo, tab, small_list, ctl);
    using MC = typename C::M;
    seg_medium_kernel<KT, V, MC::kBlock, MC::kIpt>
        <<<grid, MC::kBlock, 0, s>>>(ki, ka, ko, vi, (V*)va, vo, tab, medium_list, ctl);
  }
  switch (var) {
    case 1:
      seg_large<KT, VT, 1>(ki, ka, ko, vi, (V*)va, vo, tab, large_list, tile_start, seg_hist, seg_base, tcount, desc, ctl, L, s);
      break;
    case 2:
      seg_large<KT, VT, 2>(ki, ka, ko, vi, (V*)va, vo, tab, large_list, tile_start, seg_hist, seg_base, tcount, desc, ctl, L, s);
      break;
    default:
      seg_large<KT, VT, 0>(ki, ka, ko, vi, (V*)va, vo, tab, large_list, tile_start, seg_hist, seg_base, tcount, desc, ctl, L, s);
  }
  return check_launch();
}

template <int KT>
int seg_by_val(mgs_val vt, const void* kin, void* kout, const void* vin, void* vout, const uint64_t* offsets,
               bool ranges, const uint64_t* val_begin, size_t nseg, size_t max_total, const mgs_tuning* tuning, void* temp, size_t temp_bytes,
               hipStream_t s) {
  switch (vt) {
    case MGS_VAL_NONE:
      return seg_typed<KT, MGS_VAL_NONE>(kin, kout, vin, vout, offsets, ranges, val_begin, nseg, max_total, tuning, temp,
                                                 temp_bytes, s);
    case MGS_VAL_U32:
      return seg_typed<KT, MGS_VAL_U32>(kin, kout, vin, vout, offsets, ranges, val_begin, nseg, max_total, tuning, temp,
                                                 temp_bytes, s);
    case MGS_VAL_U64:
      return seg_typed<KT, MGS_VAL_U64>(kin, kout, vin, vout, offsets, ranges, val_begin, nseg, max_total, tuning, temp,
                                                 temp_bytes, s);
  }
  return MGS_ERR_BAD_TYPE;
}

template <int KB, int VB>
size_t seg_bytes(size_t max_total, size_t nseg, size_t max_seg, int var) {
  if (var < 0 || var >= kSegVariants) var = 0;
  return seg_layout<KB, VB>(max_total, nseg, max_seg, seg_large_tile<KB, VB>(var)).total;
}

template <int KB, int VB>
void seg_geom(int var, int32_t* block, int32_t* ipt) {
  switch (var) {
    case 1:
      *block = SegLargeGeom<KB, VB, 1>::kBlock;
      *ipt = SegLargeGeom<KB, VB, 1>::kIpt;
      break;
    case 2:
      *block = SegLargeGeom<KB, VB, 2>::kBlock;
      *ipt = SegLargeGeom<KB, VB, 2>::kIpt;
      break;
    default:
      *block = SegLargeGeom<KB, VB, 0>::kBlock;
      *ipt = SegLargeGeom<KB, VB, 0>::kIpt;
  }
}

}  // namespace

size_t segmented_temp_bytes(mgs_key kt, mgs_val vt, size_t max_total, size_t nseg, const mgs_tuning* tuning) {
  const int kb = key_bytes(kt), vb = val_bytes(vt);
  if (kb < 0 || vb < 0) return 0;
  const size_t ms = segment_bound(max_total, tuning);
  const int var = large_variant_of(tuning);
  if (kb == 4) {
    if (vb == 0) return seg_bytes<4, 0>(max_total, nseg, ms, var);
    if (vb == 4) return seg_bytes<4, 4>(max_total, nseg, ms, var);
    return seg_bytes<4, 8>(max_total, nseg, ms, var);
  }
  if (vb == 0) return seg_bytes<8, 0>(max_total, nseg, ms, var);
  if (vb == 4) return seg_bytes<8, 4>(max_total, nseg, ms, var);
  return seg_bytes<8, 8>(max_total, nseg, ms, var);
}

int segmented_variant(mgs_key kt, mgs_val vt, int var, int32_t* block, int32_t* ipt) {
  const int kb = key_bytes(kt), vb = val_bytes(vt);
  if (kb < 0 || vb < 0) return MGS_ERR_BAD_TYPE;
  if (var < 0 || var >= kSegVariants) return MGS_ERR_INVALID;
  int32_t b = 0, i = 0;
  if (kb == 4)
    (vb == 0 ? seg_geom<4, 0> : vb == 4 ? seg_geom<4, 4> : seg_geom<4, 8>)(var, &b, &i);
  else
    (vb == 0 ? seg_geom<8, 0> : vb == 4 ? seg_geom<8, 4> : seg_geom<8, 8>)(var, &b, &i);
  if (block) *block = b;
  if (ipt) *ipt = i;
  return MGS_OK;
}

int segmented_variant_count() { return kSegVariants; }

int segmented_dispatch(mgs_key kt, mgs_val vt, const void* kin, void* kout, const void* vin, void* vout,
                       const uint64_t* d_offsets, bool ranges, const uint64_t* d_val_begin, size_t nseg, size_t max_total, const mgs_tuning* tuning,
                       void* temp, size_t temp_bytes, hipStream_t s) {
  switch (kt) {
    case MGS_KEY_U32:
      return seg_by_val<MGS_KEY_U32>(vt, kin, kout, vin, vout, d_offsets, ranges, d_val_begin, nseg, max_total, tuning, temp, temp_bytes, s);
    case MGS_KEY_U64:
      return seg_by_val<MGS_KEY_U64>(vt, kin, kout, vin, vout, d_offsets, ranges, d_val_begin, nseg, max_total, tuning, temp, temp_bytes, s);
    case MGS_KEY_I32:
      return seg_by_val<MGS_KEY_I32>(vt, kin, kout, vin, vout, d_offsets, ranges, d_val_begin, nseg, max_total, tuning, temp, temp_bytes, s);
    case MGS_KEY_I64:
      return seg_by_val<MGS_KEY_I64>(vt, kin, kout, vin, vout, d_offsets, ranges, d_val_begin, nseg, max_total, tuning, temp, temp_bytes, s);
    case MGS_KEY_F32:
      return seg_by_val<MGS_KEY_F32>(vt, kin, kout, vin, vout, d_offsets, ranges, d_val_begin, nseg, max_total, tuning, temp, temp_bytes, s);
    case MGS_KEY_F64:
      return seg_by_val<MGS_KEY_F64>(vt, kin, kout, vin, vout, d_offsets, ranges, d_val_begin, nseg, max_total, tuning, temp, temp_bytes, s);
  }
  return MGS_ERR_BAD_TYPE;
}

}  // namespace mgs
