// Onesweep LSD radix sort kernels and their launch sequence (see
// radix_sort.hpp for the algorithm outline).
#include <type_traits>

#include "internal.hpp"
#include "radix_device.hpp"

namespace mgs {


// ---------------------------------------------------------------------------
// Pipeline (one call, fixed launch sequence, no host synchronisation):
//
//   init       zero histograms and tile counters
//   hist       one read: per-stripe 256-bin histograms of every byte
//              (stripes = S contiguous runs of tiles of the input)
//   plan       per byte: stripe output offsets (in place), global totals,
//              "trivial" flag (one digit holds every key)
//   msd        picks the MSD byte m = highest non-trivial byte; lays out the
//              256 buckets of byte m and their tiles
//   pass m     onesweep over byte m; each stripe is its own look-back chain
//              (the stripe counts describe the input, which this first pass
//              reads); afterwards every key sits in its byte-m bucket
//   seghist    one read: per-bucket histograms of bytes 0..m-1
//   segplan    per-bucket output offsets of bytes 0..m-1
//   sched      round-robin order of the bucket tiles (tile j of every bucket
//              before tile j+1 of any), so consecutive workgroups work on
//              different buckets and each bucket's look-back chain only
//              ever waits on the previous tile
//   pass b     b = 0..m-1 (non-trivial only): LSD onesweep inside buckets;
//              every later pass only permutes within buckets, so the bucket
//              counts stay valid
//   fixup      copies the result into keys_out if the pass count left it in
//              the scratch buffer
// ---------------------------------------------------------------------------

struct Stripes {
  uint32_t count;      // S
  uint32_t tiles_per;  // tiles per stripe (the last stripe may be shorter)
  uint32_t num_tiles;
  uint32_t tile_keys;
};

struct SortCtl {
  int32_t msd;          // MSD byte, -1 when every byte is trivial
  int32_t na;           // active passes: MSD pass + non-trivial lower bytes
  uint32_t seg_tiles;   // tiles of all bucket chains
  uint32_t seg_blocks;  // seghist workgroups
};

constexpr int kSegHistTiles = 32;  // tiles per seghist workgroup

__global__ void init_kernel(uint32_t* __restrict__ a, uint32_t a_words, uint32_t* __restrict__ b, uint32_t b_words,
                            uint32_t* __restrict__ counters) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t st = gridDim.x * blockDim.x;
  for (uint32_t i = g; i < a_words; i += st) a[i] = 0;
  for (uint32_t i = g; i < b_words; i += st) b[i] = 0;
  if (g < kMaxPasses + 1) counters[g] = 0;
}

// Per-stripe histograms of every byte.  Grid = S * G workgroups, G per
// stripe.  Also clears the look-back descriptors (desc_words 32-bit words).
// hist layout: [byte][stripe][digit]
template <int KT, int BLOCK, bool VEC>
__global__ __launch_bounds__(BLOCK) void hist_kernel(const typename Codec<KT>::Bits* __restrict__ keys, uint32_t n,
                                                     Stripes st, uint32_t per_stripe_blocks,
                                                     uint32_t* __restrict__ hist, uint32_t* __restrict__ desc,
                                                     size_t desc_words) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int P = (int)sizeof(Bits);
  __shared__ uint32_t h[P][kRadix];
  const int tid = threadIdx.x;
  for (int i = tid; i < P * kRadix; i += BLOCK) (&h[0][0])[i] = 0;
  {
    const size_t gtid = (size_t)blockIdx.x * BLOCK + tid;
    const size_t gstride = (size_t)gridDim.x * BLOCK;
    uint4* d4 = reinterpret_cast<uint4*>(desc);
    const size_t n4 = desc_words / 4;
    for (size_t i = gtid; i < n4; i += gstride) d4[i] = make_uint4(0, 0, 0, 0);
    for (size_t i = n4 * 4 + gtid; i < desc_words; i += gstride) desc[i] = 0;
  }
  __syncthreads();

  const uint32_t s = blockIdx.x / per_stripe_blocks;
  const uint32_t gb = blockIdx.x % per_stripe_blocks;
  const size_t lo = (size_t)s * st.tiles_per * st.tile_keys;
  const size_t hi = std::min<size_t>(lo + (size_t)st.tiles_per * st.tile_keys, n);
  auto count = [&](Bits k) {
#pragma unroll
    for (int p = 0; p < P; ++p) atomicAdd(&h[p][digit_of<KT>(k, 8 * p)], 1u);
  };
  const size_t t0 = (size_t)gb * BLOCK + tid;
  const size_t ts = (size_t)per_stripe_blocks * BLOCK;
  if constexpr (VEC) {
    // stripes start at multiples of the tile size: 16-byte alignment of
    // `keys` carries over to every stripe
    constexpr int V = 16 / (int)sizeof(Bits);
    using Vec = typename std::conditional<sizeof(Bits) == 4, uint4, ulonglong2>::type;
    const Vec* kv = reinterpret_cast<const Vec*>(keys + lo);
    const size_t nv = (hi - lo) / V;
    for (size_t i = t0; i < nv; i += ts) {
      Vec v = kv[i];
      if constexpr (sizeof(Bits) == 4) {
        count(v.x); count(v.y); count(v.z); count(v.w);
      } else {
        count(v.x); count(v.y);
      }
    }
    for (size_t i = lo + nv * V + t0; i < hi; i += ts) count(keys[i]);
  } else {
    for (size_t i = lo + t0; i < hi; i += ts) count(keys[i]);
  }
  __syncthreads();
  for (int i = tid; i < P * kRadix; i += BLOCK) {
    const uint32_t c = (&h[0][0])[i];
    const int p = i / kRadix, d = i % kRadix;
    if (c) atomicAdd(&hist[((size_t)p * st.count + s) * kRadix + d], c);
  }
}

// One workgroup of 1024 threads per byte: stripe output offsets (in place),
// global digit totals and the trivial flag.
__global__ __launch_bounds__(1024) void plan_kernel(uint32_t* __restrict__ hist, uint32_t n, uint32_t stripes,
                                                    uint32_t* __restrict__ glob, int32_t* __restrict__ trivial) {
  __shared__ uint32_t part[4][kRadix];
  __shared__ uint32_t base[kRadix];
  __shared__ uint32_t sums[4];
  const int p = blockIdx.x;
  const int tid = threadIdx.x;
  const int d = tid & (kRadix - 1), q = tid >> 8;  // thread (d, q): quarter q of the stripes
  uint32_t* h = hist + (size_t)p * stripes * kRadix;
  const uint32_t s0 = (stripes * q) / 4, s1 = (stripes * (q + 1)) / 4;
  uint32_t sum = 0;
#pragma unroll 8
  for (uint32_t s = s0; s < s1; ++s) sum += h[(size_t)s * kRadix + d];
  part[q][d] = sum;
  __syncthreads();
  const uint32_t total = part[0][d] + part[1][d] + part[2][d] + part[3][d];
  const int t = __syncthreads_or(q == 0 && total == n);
  const uint32_t excl = block_excl_scan_256(q == 0 ? total : 0u, sums, tid);
  if (q == 0) {
    base[d] = excl;
    glob[p * kRadix + d] = total;
  }
  __syncthreads();
  uint32_t run = base[d];
  for (int i = 0; i < q; ++i) run += part[i][d];
#pragma unroll 8
  for (uint32_t s = s0; s < s1; ++s) {
    const uint32_t c = h[(size_t)s * kRadix + d];
    h[(size_t)s * kRadix + d] = run;
    run += c;
  }
  if (tid == 0) trivial[p] = t;
}

// 256 threads: MSD byte, bucket layout, tile and seghist-block prefixes.
__global__ __launch_bounds__(256) void msd_kernel(const uint32_t* __restrict__ glob,
                                                  const int32_t* __restrict__ trivial, int passes, uint32_t tile_keys,
                                                  SortCtl* __restrict__ ctl, uint32_t* __restrict__ seg_off,
                                                  uint32_t* __restrict__ tile_start, uint32_t* __restrict__ blk_start) {
  __shared__ uint32_t sums[4];
  const int tid = threadIdx.x;
  int m = -1;
  for (int p = passes - 1; p >= 0; --p)
    if (!trivial[p]) {
      m = p;
      break;
    }
  if (m < 0) {
    if (tid == 0) *ctl = SortCtl{-1, 0, 0, 0};
    return;  // uniform branch: every thread returns
  }
  int na = 1;
  for (int p = 0; p < m; ++p) na += trivial[p] ? 0 : 1;
  const uint32_t len = glob[m * kRadix + tid];
  const uint32_t tiles = (len + tile_keys - 1) / tile_keys;
  const uint32_t blocks = (tiles + kSegHistTiles - 1) / kSegHistTiles;
  const uint32_t off = block_excl_scan_256(len, sums, tid);
  __syncthreads();
  const uint32_t tstart = block_excl_scan_256(tiles, sums, tid);
  __syncthreads();
  const uint32_t bstart = block_excl_scan_256(blocks, sums, tid);
  seg_off[tid] = off;
  tile_start[tid] = tstart;
  blk_start[tid] = bstart;
  if (tid == kRadix - 1) {
    seg_off[kRadix] = off + len;
    tile_start[kRadix] = tstart + tiles;
    blk_start[kRadix] = bstart + blocks;
    *ctl = SortCtl{m, na, tstart + tiles, bstart + blocks};
  }
}

// Place of active pass k in the ping-pong sequence.
__device__ __forceinline__ int dst_of(int k, int na, int aliased) {
  if (aliased) return (k % 2 == 0) ? (int)kBufAlt : (int)kBufOut;
  return ((na - 1 - k) % 2 == 0) ? (int)kBufOut : (int)kBufAlt;
}
__device__ __forceinline__ int src_of(int k, int na, int aliased) {
  return k == 0 ? (int)kBufIn : dst_of(k - 1, na, aliased);
}

template <typename T>
__device__ __forceinline__ const T* pick_src(int id, const T* in, const T* alt, const T* out) {
  return id == kBufIn ? in : (id == kBufAlt ? alt : out);
}
template <typename T>
__device__ __forceinline__ T* pick_dst(int id, T* alt, T* out) {
  return id == kBufAlt ? alt : out;
}

// Index of the bucket holding position x of a prefix array (last i with
// start[i] <= x), over 256 buckets.
__device__ __forceinline__ uint32_t bucket_of(const uint32_t* start, uint32_t x) {
  uint32_t lo = 0;
#pragma unroll
  for (uint32_t step = kRadix / 2; step > 0; step >>= 1)
    if (start[lo + step] <= x) lo += step;
  return lo;
}

// MSD pass over byte ctl->msd of the whole input, stripes as chains.
template <int KT, typename V, int BLOCK, int IPT, typename D>
__global__ __launch_bounds__(BLOCK) void msd_pass_kernel(
    const typename Codec<KT>::Bits* k_in, typename Codec<KT>::Bits* k_alt, typename Codec<KT>::Bits* k_out,
    const V* v_in, V* v_alt, V* v_out, uint32_t n, int aliased, Stripes st, const SortCtl* __restrict__ ctl,
    const uint32_t* __restrict__ stripe_base, uint32_t* __restrict__ tile_counter, D* __restrict__ desc) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int TILE = BLOCK * IPT;
  const int m = ctl->msd;
  if (m < 0) return;
  const int na = ctl->na;
  const int db = dst_of(0, na, aliased);
  Bits* dst = pick_dst(db, k_alt, k_out);
  V* vdst = nullptr;
  if constexpr (kHasVal<V>) vdst = pick_dst(db, v_alt, v_out);
  __shared__ uint32_t s_id;
  if (threadIdx.x == 0) {
    const uint32_t ids = st.count * st.tiles_per;
    uint32_t g, t;
    do {
      g = atomicAdd(&tile_counter[0], 1u);
      t = (g % st.count) * st.tiles_per + g / st.count;
    } while (t >= st.num_tiles && g < ids);
    s_id = g < ids ? g : 0xFFFFFFFFu;
  }
  __syncthreads();
  const uint32_t g = s_id;
  if (g == 0xFFFFFFFFu) return;
  const uint32_t s = g % st.count, j = g / st.count;
  const size_t chain0 = (size_t)s * st.tiles_per * TILE;
  const uint32_t chain_n = (uint32_t)std::min<size_t>((size_t)st.tiles_per * TILE, n - chain0);
  const V* vs = nullptr;
  if constexpr (kHasVal<V>) vs = v_in + chain0;
  onesweep_tile<KT, V, BLOCK, IPT, D>(k_in + chain0, dst, vs, vdst, chain_n, n, j,
                                      desc + (size_t)s * st.tiles_per * kRadix,
                                      stripe_base + ((size_t)m * st.count + s) * kRadix, m * 8, 0);
}

// Per-bucket histograms of bytes 0..m-1 of the MSD pass output.
// seg_hist layout: [bucket][byte][digit]
template <int KT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void seghist_kernel(const typename Codec<KT>::Bits* k_in,
                                                        const typename Codec<KT>::Bits* k_alt,
                                                        const typename Codec<KT>::Bits* k_out, int aliased,
                                                        uint32_t tile_keys, const SortCtl* __restrict__ ctl,
                                                        const uint32_t* __restrict__ seg_off,
                                                        const uint32_t* __restrict__ blk_start,
                                                        uint32_t* __restrict__ seg_hist) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int P = (int)sizeof(Bits);
  __shared__ uint32_t h[P - 1][kRadix];
  __shared__ uint32_t sb[kRadix + 1];
  const int m = ctl->msd;
  if (m <= 0 || blockIdx.x >= ctl->seg_blocks) return;
  const int tid = threadIdx.x;
  for (int i = tid; i < (P - 1) * kRadix; i += BLOCK) (&h[0][0])[i] = 0;
  for (int i = tid; i <= kRadix; i += BLOCK) sb[i] = blk_start[i];
  __syncthreads();
  const Bits* src = pick_src(dst_of(0, ctl->na, aliased), k_in, k_alt, k_out);
  const uint32_t c = bucket_of(sb, blockIdx.x);
  const uint32_t chunk = (uint32_t)kSegHistTiles * tile_keys;
  const uint32_t b0 = seg_off[c] + (blockIdx.x - sb[c]) * chunk;
  const uint32_t b1 = min(b0 + chunk, seg_off[c + 1]);
  for (uint32_t i = b0 + tid; i < b1; i += BLOCK) {
    const Bits k = src[i];
    for (int p = 0; p < m; ++p) atomicAdd(&h[p][digit_of<KT>(k, 8 * p)], 1u);
  }
  __syncthreads();
  for (int i = tid; i < m * kRadix; i += BLOCK) {
    const uint32_t v = (&h[0][0])[i];
    if (v) atomicAdd(&seg_hist[((size_t)c * (P - 1)) * kRadix + i], v);
  }
}

// One workgroup per bucket: output offsets of bytes 0..m-1, in place.
template <int P>
__global__ __launch_bounds__(256) void segplan_kernel(uint32_t* seg_hist, const SortCtl* __restrict__ ctl,
                                                      const uint32_t* __restrict__ seg_off) {
  __shared__ uint32_t sums[4];
  const int m = ctl->msd;
  const uint32_t c = blockIdx.x;
  const int tid = threadIdx.x;
  const uint32_t base = seg_off[c];
  for (int p = 0; p < m; ++p) {
    uint32_t* row = seg_hist + ((size_t)c * (P - 1) + p) * kRadix;
    const uint32_t v = row[tid];
    const uint32_t e = block_excl_scan_256(v, sums, tid);
    __syncthreads();
    row[tid] = base + e;
  }
}

// Round-robin schedule of the bucket tiles: position of tile j of bucket c =
// sum over buckets of min(j, tiles) + number of lower buckets with > j tiles.
__global__ __launch_bounds__(256) void sched_kernel(const SortCtl* __restrict__ ctl,
                                                    const uint32_t* __restrict__ tile_start,
                                                    uint32_t* __restrict__ sched) {
  __shared__ uint32_t ts[kRadix + 1];
  __shared__ uint32_t tc[kRadix];
  if (ctl->msd <= 0) return;
  const int tid = threadIdx.x;
  for (int i = tid; i <= kRadix; i += blockDim.x) ts[i] = tile_start[i];
  __syncthreads();
  for (int i = tid; i < kRadix; i += blockDim.x) tc[i] = ts[i + 1] - ts[i];
  __syncthreads();
  const uint32_t t = blockIdx.x * blockDim.x + tid;
  if (t >= ctl->seg_tiles) return;
  const uint32_t c = bucket_of(ts, t);
  const uint32_t j = t - ts[c];
  uint32_t pos = 0;
  for (uint32_t i = 0; i < kRadix; ++i) {
    const uint32_t ti = tc[i];
    pos += min(j, ti) + ((i < c && ti > j) ? 1u : 0u);
  }
  sched[pos] = t;
}

// LSD pass over byte `pass` (< msd) inside the buckets.
template <int KT, typename V, int BLOCK, int IPT, typename D>
__global__ __launch_bounds__(BLOCK) void lsd_pass_kernel(
    const typename Codec<KT>::Bits* k_in, typename Codec<KT>::Bits* k_alt, typename Codec<KT>::Bits* k_out,
    const V* v_in, V* v_alt, V* v_out, uint32_t n, int pass, int aliased, const SortCtl* __restrict__ ctl,
    const int32_t* __restrict__ trivial, const uint32_t* __restrict__ seg_off,
    const uint32_t* __restrict__ tile_start, const uint32_t* __restrict__ sched,
    const uint32_t* __restrict__ seg_base, uint32_t* __restrict__ tile_counter, D* __restrict__ desc) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int P = (int)sizeof(Bits);
  const int m = ctl->msd;
  if (pass >= m || trivial[pass]) return;
  int kb = 0;
  for (int p = 0; p < pass; ++p) kb += trivial[p] ? 0 : 1;
  const int k = 1 + kb, na = ctl->na;
  const int sb = src_of(k, na, aliased), db = dst_of(k, na, aliased);
  const Bits* src = pick_src(sb, k_in, (const Bits*)k_alt, (const Bits*)k_out);
  Bits* dst = pick_dst(db, k_alt, k_out);
  const V* vsrc = nullptr;
  V* vdst = nullptr;
  if constexpr (kHasVal<V>) {
    vsrc = pick_src(sb, v_in, (const V*)v_alt, (const V*)v_out);
    vdst = pick_dst(db, v_alt, v_out);
  }
  __shared__ uint32_t s_t;
  if (threadIdx.x == 0) {
    const uint32_t g = atomicAdd(&tile_counter[1 + pass], 1u);
    s_t = g < ctl->seg_tiles ? sched[g] : 0xFFFFFFFFu;
  }
  __syncthreads();
  const uint32_t t = s_t;
  if (t == 0xFFFFFFFFu) return;
  const uint32_t c = bucket_of(tile_start, t);
  const uint32_t j = t - tile_start[c];
  const uint32_t b0 = seg_off[c], len = seg_off[c + 1] - b0;
  const V* vs = nullptr;
  if constexpr (kHasVal<V>) vs = vsrc + b0;
  onesweep_tile<KT, V, BLOCK, IPT, D>(src + b0, dst, vs, vdst, len, n, j, desc + (size_t)tile_start[c] * kRadix,
                                      seg_base + ((size_t)c * (P - 1) + pass) * kRadix, pass * 8, kb & 1);
}

// Copies the result into the output buffers when the passes left it in the
// scratch buffer (or did not run).
template <typename Bits, typename V>
__global__ void fixup_kernel(const Bits* k_in, const Bits* k_alt, Bits* k_out, const V* v_in, const V* v_alt, V* v_out,
                             uint32_t n, const SortCtl* __restrict__ ctl, int aliased) {
  const int na = ctl->na;
  int s;
  if (na == 0) {
    s = kBufIn;
  } else {
    s = dst_of(na - 1, na, aliased);
    if (s == kBufOut) return;
  }
  const Bits* ks = s == kBufIn ? k_in : k_alt;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stp = (size_t)gridDim.x * blockDim.x;
  if (ks != k_out)
    for (size_t i = g; i < n; i += stp) k_out[i] = ks[i];
  if constexpr (kHasVal<V>) {
    const V* vs = s == kBufIn ? v_in : v_alt;
    if (vs != v_out)
      for (size_t i = g; i < n; i += stp) v_out[i] = vs[i];
  }
}

// ---------------------------------------------------------------------------
// Single-workgroup sort of up to one tile, entirely in LDS.
// ---------------------------------------------------------------------------
template <int KT, typename V, int BLOCK, int IPT>
__global__ __launch_bounds__(BLOCK) void small_sort_kernel(const typename Codec<KT>::Bits* k_in,
                                                           typename Codec<KT>::Bits* k_out, const V* v_in, V* v_out,
                                                           uint32_t n) {
  small_sort_block<KT, V, BLOCK, IPT>(k_in, k_out, v_in, v_out, n);
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

#ifndef MGS_STRIPES
#define MGS_STRIPES 256
#endif

struct SortLayout {
  size_t alt_k, alt_v, hist, glob, trivial, ctl, seg_off, tile_start, blk_start, seg_hist, sched, counters, desc_a,
      desc_b, total;
  Stripes st;
  uint32_t seg_tiles_ub;
  bool wide_desc;
};

template <int KB, int VB>
SortLayout sort_layout(size_t n) {
  using Cfg = TileCfg<KB, VB>;
  constexpr int P = KB;
  SortLayout L{};
  L.st.tile_keys = Cfg::kTile;
  L.st.num_tiles = (uint32_t)((n + Cfg::kTile - 1) / Cfg::kTile);
  L.st.tiles_per = (L.st.num_tiles + MGS_STRIPES - 1) / MGS_STRIPES;
  L.st.count = (L.st.num_tiles + L.st.tiles_per - 1) / L.st.tiles_per;
  L.seg_tiles_ub = L.st.num_tiles + kRadix;
  L.wide_desc = n > (size_t(1) << 30);
  const size_t dw = L.wide_desc ? 8 : 4;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align_up(o + bytes);
    return at;
  };
  L.alt_k = take(n * KB);
  L.alt_v = take(n * VB);
  L.hist = take((size_t)P * L.st.count * kRadix * 4);
  L.glob = take((size_t)P * kRadix * 4);
  L.trivial = take(kMaxPasses * 4);
  L.ctl = take(sizeof(SortCtl));
  L.seg_off = take((kRadix + 1) * 4);
  L.tile_start = take((kRadix + 1) * 4);
  L.blk_start = take((kRadix + 1) * 4);
  L.seg_hist = take((size_t)kRadix * (P - 1) * kRadix * 4);
  L.sched = take((size_t)L.seg_tiles_ub * 4);
  L.counters = take((kMaxPasses + 1) * 4);
  L.desc_a = take((size_t)L.st.num_tiles * kRadix * dw);  // must stay adjacent:
  L.desc_b = take((size_t)L.seg_tiles_ub * kRadix * dw);  // cleared as one range
  L.total = o;
  return L;
}

template <int KT, int VT>
int sort_typed(const void* kin, void* kout, const void* vin, void* vout, size_t n, void* temp, size_t temp_bytes,
               hipStream_t s) {
  using Bits = typename Codec<KT>::Bits;
  using V = typename ValOf<VT>::T;
  constexpr int KB = sizeof(Bits);
  constexpr int VB = kValBytes<V>;
  constexpr int P = KB;
  using Cfg = TileCfg<KB, VB>;
  using SCfg = SmallCfg<KB, VB>;
  using VP = typename std::conditional<kHasVal<V>, V, uint32_t>::type;

  const Bits* ki = static_cast<const Bits*>(kin);
  Bits* ko = static_cast<Bits*>(kout);
  const V* vi = static_cast<const V*>(vin);
  V* vo = static_cast<V*>(vout);

  if (n <= (size_t)SCfg::kTile) {
    small_sort_kernel<KT, V, SCfg::kBlock, SCfg::kIpt><<<1, SCfg::kBlock, 0, s>>>(ki, ko, vi, vo, (uint32_t)n);
    return check_launch();
  }

  const SortLayout L = sort_layout<KB, VB>(n);
  if (temp == nullptr || temp_bytes < L.total) return MGS_ERR_TEMP_TOO_SMALL;
  char* t = static_cast<char*>(temp);
  Bits* ka = reinterpret_cast<Bits*>(t + L.alt_k);
  V* va = reinterpret_cast<V*>(reinterpret_cast<VP*>(t + L.alt_v));
  uint32_t* hist = reinterpret_cast<uint32_t*>(t + L.hist);
  uint32_t* glob = reinterpret_cast<uint32_t*>(t + L.glob);
  int32_t* trivial = reinterpret_cast<int32_t*>(t + L.trivial);
  SortCtl* ctl = reinterpret_cast<SortCtl*>(t + L.ctl);
  uint32_t* seg_off = reinterpret_cast<uint32_t*>(t + L.seg_off);
  uint32_t* tile_start = reinterpret_cast<uint32_t*>(t + L.tile_start);
  uint32_t* blk_start = reinterpret_cast<uint32_t*>(t + L.blk_start);
  uint32_t* seg_hist = reinterpret_cast<uint32_t*>(t + L.seg_hist);
  uint32_t* sched = reinterpret_cast<uint32_t*>(t + L.sched);
  uint32_t* counters = reinterpret_cast<uint32_t*>(t + L.counters);
  void* desc_a = t + L.desc_a;
  void* desc_b = t + L.desc_b;
  const size_t desc_words = (L.total - L.desc_a) / 4;
  const int aliased = (kin == kout) || (kHasVal<V> && vin == vout);
  const uint32_t hist_words = (uint32_t)(P * L.st.count * kRadix);
  const uint32_t seg_words = (uint32_t)(kRadix * (P - 1) * kRadix);

  init_kernel<<<64, 1024, 0, s>>>(hist, hist_words, seg_hist, seg_words, counters);
  {
    constexpr int HB = 256;
    const size_t stripe_keys = (size_t)L.st.tiles_per * L.st.tile_keys;
    const uint32_t g = (uint32_t)std::max<size_t>(
        1, std::min<size_t>(2048 / L.st.count, (stripe_keys + 16 * 1024 - 1) / (16 * 1024)));
    const unsigned grid = L.st.count * g;
    if ((reinterpret_cast<uintptr_t>(kin) % 16) == 0)
      hist_kernel<KT, HB, true><<<grid, HB, 0, s>>>(ki, (uint32_t)n, L.st, g, hist, (uint32_t*)desc_a, desc_words);
    else
      hist_kernel<KT, HB, false><<<grid, HB, 0, s>>>(ki, (uint32_t)n, L.st, g, hist, (uint32_t*)desc_a, desc_words);
  }
  plan_kernel<<<P, 1024, 0, s>>>(hist, (uint32_t)n, L.st.count, glob, trivial);
  msd_kernel<<<1, 256, 0, s>>>(glob, trivial, P, Cfg::kTile, ctl, seg_off, tile_start, blk_start);
  if (L.wide_desc)
    msd_pass_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, uint64_t><<<L.st.num_tiles, Cfg::kBlock, 0, s>>>(
        ki, ka, ko, vi, va, vo, (uint32_t)n, aliased, L.st, ctl, hist, counters, (uint64_t*)desc_a);
  else
    msd_pass_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, uint32_t><<<L.st.num_tiles, Cfg::kBlock, 0, s>>>(
        ki, ka, ko, vi, va, vo, (uint32_t)n, aliased, L.st, ctl, hist, counters, (uint32_t*)desc_a);
  {
    const unsigned grid = (unsigned)(L.seg_tiles_ub / kSegHistTiles + kRadix + 1);
    seghist_kernel<KT, 256><<<grid, 256, 0, s>>>(ki, ka, ko, aliased, Cfg::kTile, ctl, seg_off, blk_start, seg_hist);
  }
  segplan_kernel<P><<<kRadix, 256, 0, s>>>(seg_hist, ctl, seg_off);
  sched_kernel<<<(L.seg_tiles_ub + 255) / 256, 256, 0, s>>>(ctl, tile_start, sched);
  for (int p = 0; p < P - 1; ++p) {
    if (L.wide_desc)
      lsd_pass_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, uint64_t><<<L.seg_tiles_ub, Cfg::kBlock, 0, s>>>(
          ki, ka, ko, vi, va, vo, (uint32_t)n, p, aliased, ctl, trivial, seg_off, tile_start, sched, seg_hist,
          counters, (uint64_t*)desc_b);
    else
      lsd_pass_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, uint32_t><<<L.seg_tiles_ub, Cfg::kBlock, 0, s>>>(
          ki, ka, ko, vi, va, vo, (uint32_t)n, p, aliased, ctl, trivial, seg_off, tile_start, sched, seg_hist,
          counters, (uint32_t*)desc_b);
  }
  {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    fixup_kernel<Bits, V><<<grid, 256, 0, s>>>(ki, ka, ko, vi, va, vo, (uint32_t)n, ctl, aliased);
  }
  return check_launch();
}

template <int KT>
int sort_by_val(mgs_val vt, const void* kin, void* kout, const void* vin, void* vout, size_t n, void* temp,
                size_t temp_bytes, hipStream_t s) {
  switch (vt) {
    case MGS_VAL_NONE: return sort_typed<KT, MGS_VAL_NONE>(kin, kout, vin, vout, n, temp, temp_bytes, s);
    case MGS_VAL_U32: return sort_typed<KT, MGS_VAL_U32>(kin, kout, vin, vout, n, temp, temp_bytes, s);
    case MGS_VAL_U64: return sort_typed<KT, MGS_VAL_U64>(kin, kout, vin, vout, n, temp, temp_bytes, s);
  }
  return MGS_ERR_BAD_TYPE;
}

}  // namespace

size_t sort_temp_bytes(mgs_key kt, mgs_val vt, size_t n) {
  const int kb = key_bytes(kt), vb = val_bytes(vt);
  if (kb <= 0 || vb < 0) return 0;
  size_t small_tile;
  if (kb == 4 && vb == 0) small_tile = SmallCfg<4, 0>::kTile;
  else if (kb + vb <= 8) small_tile = SmallCfg<4, 4>::kTile;
  else small_tile = SmallCfg<8, 8>::kTile;
  if (n <= small_tile) return 0;
  if (kb == 4) {
    if (vb == 0) return sort_layout<4, 0>(n).total;
    if (vb == 4) return sort_layout<4, 4>(n).total;
    return sort_layout<4, 8>(n).total;
  }
  if (vb == 0) return sort_layout<8, 0>(n).total;
  if (vb == 4) return sort_layout<8, 4>(n).total;
  return sort_layout<8, 8>(n).total;
}

int sort_dispatch(mgs_key kt, mgs_val vt, const void* kin, void* kout, const void* vin, void* vout, size_t n,
                  void* temp, size_t temp_bytes, hipStream_t s) {
  switch (kt) {
    case MGS_KEY_U32: return sort_by_val<MGS_KEY_U32>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s);
    case MGS_KEY_U64: return sort_by_val<MGS_KEY_U64>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s);
    case MGS_KEY_I32: return sort_by_val<MGS_KEY_I32>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s);
    case MGS_KEY_I64: return sort_by_val<MGS_KEY_I64>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s);
    case MGS_KEY_F32: return sort_by_val<MGS_KEY_F32>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s);
    case MGS_KEY_F64: return sort_by_val<MGS_KEY_F64>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s);
  }
  return MGS_ERR_BAD_TYPE;
}

}  // namespace mgs
