// Three-pass hybrid radix sort for large arrays of 4-byte keys (u32, i32,
// f32; keys alone), the path of mgs_sort for the in-HBM sorts of 2^26 keys
// and more.  The plain LSD sort (radix_sort.hip) reads and writes the array
// four times; this one does it three times:
//
//   upfront  one read: every workgroup counts the top 16 bits of the encoded
//            keys of its range (its pass-1 chain) in a 65536-bin LDS
//            histogram of 16-bit counters, and stores it as its slice
//   scan     per top byte b3: the sub-bucket sizes (b3, b2) summed over the
//            slices, their offsets inside bucket b3, the chains' prefix of
//            every b3 digit, and the batches of the local sort; a bucket or
//            sub-bucket too large for the local sort, or a 16-bit counter
//            that overflowed, turns the path off (the LSD sort runs instead)
//   finish   one workgroup: bucket starts, batch list offsets, tile counters
//   pass 1   MSD scatter by byte 3: workgroup c streams chain c (a fixed
//            range of the input) tile by tile with running per-digit output
//            offsets -- the chain's digit bases come from the upfront
//            histograms, so there is no look-back between workgroups
//   pass 2   scatter by byte 2 inside each byte-3 bucket: workgroup b streams
//            bucket b the same way, bases from the sub-bucket offsets
//   local    every batch of consecutive sub-buckets (at most kLocalCap keys,
//            one top-16-bit value, or several of one bucket) is read once,
//            sorted by its low bytes in LDS and written back in place
//
// Load balance: a workgroup that runs out of work takes tiles from the back
// end of another chain (at most one such helper per chain); a back-end
// worker places each tile just below the chain's keys already placed from
// the back, using the chain's digit totals, so front and back workers never
// wait on each other.  All keys stay stable: pass 1 keeps input order inside
// each digit, pass 2 keeps pass-1 order, and the local sort is an LSD sort
// of the batch.
//
// Every kernel of the path checks the decision word first; when the path is
// off, they return at once and the LSD sort, launched after them, runs.
#include <cstdlib>
#include <map>
#include <mutex>

#include "internal.hpp"
#include "msd_sort.hpp"
#include "radix_device.hpp"

namespace mgs {

namespace {

constexpr int kMsdBlock = MsdCfg::kBlock;
constexpr int kMsdIpt = MsdCfg::kIpt;
constexpr int kMsdTile = MsdCfg::kTile;
constexpr int kLocalIpt = MsdCfg::kLocalIpt;
constexpr uint32_t kLocalCap = MsdCfg::kLocalCap;
constexpr uint32_t kBatchTarget = kLocalCap / 2;
constexpr uint32_t kNoWork = 0xFFFFFFFFu;

// Decision word bits (MsdCtl::off): any set bit turns the path off.
enum : uint32_t { kOffOverflow = 1u, kOffSubBucket = 2u, kOffBucket = 4u };

__device__ __forceinline__ uint32_t top16(uint32_t enc) { return enc >> 16; }

// ---------------------------------------------------------------------------
// upfront
// ---------------------------------------------------------------------------
template <int KT, bool VEC>
__global__ __launch_bounds__(kMsdBlock) void msd_upfront_kernel(const uint32_t* __restrict__ keys, uint32_t n,
                                                               uint32_t chain_keys, MsdCtl* __restrict__ ctl,
                                                               uint32_t* __restrict__ slices,
                                                               uint32_t* __restrict__ chain_hist,
                                                               uint32_t* __restrict__ chain_bad,
                                                               uint32_t* __restrict__ status) {
  // 65536 16-bit counters, two per word: bin b in word b >> 1, half b & 1
  __shared__ uint32_t h[kMsdBins / 2];
  __shared__ uint32_t s_rows[kRadix];
  const int tid = threadIdx.x;
  const uint32_t c = blockIdx.x;
  if (c == 0 && tid == 0) {  // nothing else writes these during this kernel
    *status = 0;
    ctl->off = 0;
  }
  for (int i = tid; i < kMsdBins / 2; i += kMsdBlock) h[i] = 0;
  __syncthreads();
  const size_t lo = (size_t)c * chain_keys;
  const size_t hi = std::min<size_t>((size_t)n, lo + chain_keys);
  auto count = [&](uint32_t k) {
    const uint32_t b = top16(Codec<KT>::enc(k));
    __hip_atomic_fetch_add(&h[b >> 1], 1u << ((b & 1u) << 4), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  if (lo < hi) {
    if constexpr (VEC) {
      // chain_keys is a multiple of 4, so [lo, hi) starts 16-byte aligned
      constexpr int U = 4;
      const uint4* kv = reinterpret_cast<const uint4*>(keys + lo);
      const size_t nv = (hi - lo) / 4;
      size_t i = tid;
      for (; i + (U - 1) * kMsdBlock < nv; i += U * kMsdBlock) {
        uint4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = kv[i + u * kMsdBlock];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          count(v[u].x);
          count(v[u].y);
          count(v[u].z);
          count(v[u].w);
        }
      }
      for (; i < nv; i += kMsdBlock) {
        const uint4 v = kv[i];
        count(v.x);
        count(v.y);
        count(v.z);
        count(v.w);
      }
      for (size_t k = lo + nv * 4 + tid; k < hi; k += kMsdBlock) count(keys[k]);
    } else {
      for (size_t k = lo + tid; k < hi; k += kMsdBlock) count(keys[k]);
    }
  }
  __syncthreads();
  // byte-3 histogram of the chain: row r is the 128 words [128 r, 128 r + 128)
  const int lane = tid & (kWave - 1), w = tid / kWave;
  for (int r = w; r < kRadix; r += kMsdBlock / kWave) {
    const uint32_t a = h[r * 128 + lane], b = h[r * 128 + 64 + lane];
    uint32_t s = (a & 0xFFFFu) + (a >> 16) + (b & 0xFFFFu) + (b >> 16);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    if (lane == 0) s_rows[r] = s;
  }
  __syncthreads();
  if (tid < kRadix) chain_hist[(size_t)c * kRadix + tid] = s_rows[tid];
  if (tid < kWave) {
    // a wrapped 16-bit counter loses 65536 (high half) or 65535 (low half,
    // whose carry lands in the high half): the rows then fall short of the
    // range, and the path is turned off
    uint32_t s = s_rows[tid] + s_rows[tid + 64] + s_rows[tid + 128] + s_rows[tid + 192];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    if (tid == 0) chain_bad[c] = s != (uint32_t)(hi > lo ? hi - lo : 0) ? 1u : 0u;
  }
  uint4* dst = reinterpret_cast<uint4*>(slices + (size_t)c * (kMsdBins / 2));
  for (int i = tid; i < kMsdBins / 8; i += kMsdBlock)
    dst[i] = make_uint4(h[4 * i], h[4 * i + 1], h[4 * i + 2], h[4 * i + 3]);
}

// ---------------------------------------------------------------------------
// scan: one workgroup of 256 threads per byte-3 bucket b, thread d = byte 2
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void msd_scan_kernel(const uint32_t* __restrict__ slices, uint32_t chains, uint32_t n,
                                                       const uint32_t* __restrict__ chain_hist,
                                                       uint32_t* __restrict__ chain_pref, MsdCtl* __restrict__ ctl,
                                                       uint32_t* __restrict__ sub_cnt, uint32_t* __restrict__ sub_off,
                                                       uint32_t* __restrict__ bucket_tot,
                                                       uint32_t* __restrict__ leaders,
                                                       uint32_t* __restrict__ batch_cnt,
                                                       const uint32_t* __restrict__ chain_bad) {
  __shared__ uint32_t sums[4];
  __shared__ uint32_t s_carry;
  const uint32_t b = blockIdx.x, d = threadIdx.x;
  const uint32_t s = b * kRadix + d;
  if (b == 0)
    for (uint32_t c = d; c < chains; c += kRadix)
      if (chain_bad[c]) atomicOr(&ctl->off, kOffOverflow);
  // sub-bucket size: the chains' 16-bit counters of bin s
  uint32_t cnt = 0;
  {
    const uint32_t word = s >> 1, sh = (s & 1u) << 4;
    uint32_t c = 0;
    for (; c + 4 <= chains; c += 4) {
      const uint32_t a0 = slices[(size_t)c * (kMsdBins / 2) + word];
      const uint32_t a1 = slices[(size_t)(c + 1) * (kMsdBins / 2) + word];
      const uint32_t a2 = slices[(size_t)(c + 2) * (kMsdBins / 2) + word];
      const uint32_t a3 = slices[(size_t)(c + 3) * (kMsdBins / 2) + word];
      cnt += ((a0 >> sh) & 0xFFFFu) + ((a1 >> sh) & 0xFFFFu) + ((a2 >> sh) & 0xFFFFu) + ((a3 >> sh) & 0xFFFFu);
    }
    for (; c < chains; ++c) cnt += (slices[(size_t)c * (kMsdBins / 2) + word] >> sh) & 0xFFFFu;
  }
  sub_cnt[s] = cnt;
  if (cnt > kLocalCap) atomicOr(&ctl->off, kOffSubBucket);
  const uint32_t excl = block_excl_scan_256(cnt, sums, (int)d);
  sub_off[s] = excl;
  if (d == kRadix - 1) {
    bucket_tot[b] = excl + cnt;
    // a bucket far above the mean would leave pass 2 to one or two
    // workgroups: the LSD sort, whose tiles any workgroup takes, does better
    if (excl + cnt > 2 * (n / kRadix) + 2 * (uint32_t)kMsdTile) atomicOr(&ctl->off, kOffBucket);
  }

  // batches of the local sort inside this bucket: a sub-bucket starts a new
  // batch when it is the first, when it or its predecessor holds more than
  // kBatchTarget keys, or when floor(offset / kBatchTarget) changes; every
  // batch then holds fewer than 2 kBatchTarget = kLocalCap keys
  __shared__ uint32_t s_prev[kRadix];
  s_prev[d] = cnt;
  __syncthreads();
  const uint32_t pcnt = d ? s_prev[d - 1] : 0u;
  __syncthreads();
  s_prev[d] = excl;
  __syncthreads();
  const uint32_t poff = d ? s_prev[d - 1] : 0u;
  const bool lead = d == 0 || cnt > kBatchTarget || pcnt > kBatchTarget || excl / kBatchTarget != poff / kBatchTarget;
  __syncthreads();
  const uint32_t rank = block_excl_scan_256(lead ? 1u : 0u, sums, (int)d);
  if (lead) leaders[s - d + rank] = d;
  if (d == kRadix - 1) batch_cnt[b] = rank + (lead ? 1u : 0u);

  // chain prefixes of digit b for pass 1: keys with byte 3 == b in chains
  // before chain c
  if (d == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < chains; c0 += kRadix) {
    const uint32_t c = c0 + d;
    const uint32_t v = c < chains ? chain_hist[(size_t)c * kRadix + b] : 0u;
    const uint32_t e = block_excl_scan_256(v, sums, (int)d);
    const uint32_t carry = s_carry;
    if (c < chains) chain_pref[(size_t)c * kRadix + b] = carry + e;
    __syncthreads();
    if (d == kRadix - 1) s_carry = carry + e + v;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// finish: one workgroup of 256 threads
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void msd_finish_kernel(MsdCtl* __restrict__ ctl,
                                                         const uint32_t* __restrict__ bucket_tot,
                                                         uint32_t* __restrict__ bucket_start,
                                                         const uint32_t* __restrict__ batch_cnt,
                                                         uint32_t* __restrict__ batch_base, uint32_t n,
                                                         uint32_t chain_keys, uint32_t chains,
                                                         uint64_t* __restrict__ ctr1, uint64_t* __restrict__ ctr2,
                                                         uint32_t* __restrict__ claim1, uint32_t* __restrict__ claim2) {
  __shared__ uint32_t sums[4];
  const uint32_t d = threadIdx.x;
  const uint32_t bt = bucket_tot[d];
  const uint32_t bs = block_excl_scan_256(bt, sums, (int)d);
  bucket_start[d] = bs;
  if (d == kRadix - 1) bucket_start[kRadix] = bs + bt;
  __syncthreads();  // every thread has read the scan's partial sums
  const uint32_t bc = batch_cnt[d];
  const uint32_t bb = block_excl_scan_256(bc, sums, (int)d);
  batch_base[d] = bb;
  if (d == kRadix - 1) {
    batch_base[kRadix] = bb + bc;
    ctl->batches = bb + bc;
    ctl->batch_next = 0;
  }
  // tile counters: front index in the low word, back index (the tile count)
  // in the high word
  ctr2[d] = (uint64_t)((bt + kMsdTile - 1) / kMsdTile) << 32;
  claim2[d] = 0;
  for (uint32_t c = d; c < chains; c += kRadix) {
    const size_t lo = (size_t)c * chain_keys;
    const uint32_t len = lo < n ? (uint32_t)std::min<size_t>(chain_keys, n - lo) : 0u;
    ctr1[c] = (uint64_t)((len + kMsdTile - 1) / kMsdTile) << 32;
    claim1[c] = 0;
  }
}

// ---------------------------------------------------------------------------
// passes 1 and 2
// ---------------------------------------------------------------------------
struct MsdPassArgs {
  const uint32_t* src;
  uint32_t* dst;
  uint32_t n;
  uint32_t chains;      // pass 1: input chains; pass 2: 256 buckets
  uint32_t chain_keys;  // pass 1: keys per chain
  int shift;            // 24 or 16
  const uint32_t* bucket_start;
  const uint32_t* chain_hist;  // pass 1 digit totals per chain
  const uint32_t* chain_pref;  // pass 1 prefix over chains per digit
  const uint32_t* sub_cnt;     // pass 2 digit totals per bucket
  const uint32_t* sub_off;     // pass 2 digit offsets inside each bucket
  uint64_t* ctr;
  uint32_t* claim;
  const MsdCtl* ctl;
};

// Takes the next tile of chain c from the front or the back end.
__device__ __forceinline__ uint32_t msd_take(uint64_t* ctr, uint32_t c, bool back) {
  const uint64_t old = atomicAdd(reinterpret_cast<unsigned long long*>(&ctr[c]),
                                 back ? 0xFFFFFFFF00000000ull : 1ull);
  const uint32_t f = (uint32_t)old, b = (uint32_t)(old >> 32);
  if (f >= b) return kNoWork;
  return back ? b - 1 : f;
}

// Claims the back end of the first chain, from `start` on (cyclically), that
// has tiles left and no back-end worker yet, and takes its last free tile.
__device__ __forceinline__ void msd_find_back(uint64_t* ctr, uint32_t* claim, uint32_t chains, uint32_t start,
                                              uint32_t& chain, uint32_t& tile) {
  chain = kNoWork;
  tile = kNoWork;
  for (uint32_t i = 0; i < chains; ++i) {
    const uint32_t c = (start + i) % chains;
    const uint64_t v = __hip_atomic_load(&ctr[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)v >= (uint32_t)(v >> 32)) continue;
    if (__hip_atomic_load(&claim[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) continue;
    if (atomicCAS(&claim[c], 0u, 1u) != 0u) continue;
    const uint32_t t = msd_take(ctr, c, true);
    if (t != kNoWork) {
      chain = c;
      tile = t;
      return;
    }
  }
}

template <int KT, int PASS>
__global__ __launch_bounds__(kMsdBlock, 4) void msd_pass_kernel(MsdPassArgs a) {
  constexpr int W = kMsdBlock / kWave;
  constexpr int IPT = kMsdIpt;
  constexpr int TILE = kMsdTile;
  if (a.ctl->off) return;

  __shared__ uint32_t keys[TILE];
  __shared__ uint32_t off[W][kRadix];
  __shared__ uint32_t run[kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint32_t sums[4];
  __shared__ uint32_t s_tile, s_chain, s_back;

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;
  for (int i = tid; i < W * kRadix; i += kMsdBlock) (&off[0][0])[i] = 0;

  // chain bounds and digit bases
  auto chain_lo = [&](uint32_t c) -> uint32_t {
    if constexpr (PASS == 1) return c * a.chain_keys;
    else return a.bucket_start[c];
  };
  auto chain_hi = [&](uint32_t c) -> uint32_t {
    if constexpr (PASS == 1) return (uint32_t)std::min<size_t>((size_t)(c + 1) * a.chain_keys, a.n);
    else return a.bucket_start[c + 1];
  };
  // thread d < 256: first output position of digit d of chain c (front) or
  // one past its last (back)
  auto digit_start = [&](uint32_t c, uint32_t d, bool back) -> uint32_t {
    uint32_t base, tot;
    if constexpr (PASS == 1) {
      base = a.bucket_start[d] + a.chain_pref[(size_t)c * kRadix + d];
      tot = a.chain_hist[(size_t)c * kRadix + d];
    } else {
      base = a.bucket_start[c] + a.sub_off[c * kRadix + d];
      tot = a.sub_cnt[c * kRadix + d];
    }
    return back ? base + tot : base;
  };

  // the workgroup's first unit of work: its own chain from the front, else
  // the back end of another chain
  if (tid == 0) {
    uint32_t c = kNoWork, t = kNoWork, bk = 0;
    if (blockIdx.x < a.chains) {
      c = blockIdx.x;
      t = msd_take(a.ctr, c, false);
    }
    if (t == kNoWork) {
      msd_find_back(a.ctr, a.claim, a.chains, blockIdx.x, c, t);
      bk = 1;
    }
    s_chain = c;
    s_tile = t;
    s_back = bk;
  }
  __syncthreads();
  if (s_tile == kNoWork) return;
  if (tid < kRadix) run[tid] = digit_start(s_chain, (uint32_t)tid, s_back != 0);
  __syncthreads();

  uint32_t chain = s_chain, tile = s_tile;
  bool back = s_back != 0;
  uint32_t clo = chain_lo(chain), chi = chain_hi(chain);
  uint32_t kc[IPT];
  auto load = [&](uint32_t lo, uint32_t hi, uint32_t t, uint32_t (&k)[IPT]) {
    const uint32_t tb = lo + t * (uint32_t)TILE;
    const uint32_t wb = tb + (uint32_t)(w * IPT * kWave + lane);
    if (hi - tb >= (uint32_t)TILE) {
#pragma unroll
      for (int i = 0; i < IPT; ++i) k[i] = a.src[wb + i * kWave];
    } else {
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint32_t idx = wb + i * kWave;
        k[i] = idx < hi ? a.src[idx] : 0u;
      }
    }
  };
  load(clo, chi, tile, kc);

  for (;;) {
    const uint32_t tb = clo + tile * (uint32_t)TILE;
    const uint32_t tile_cnt = min((uint32_t)TILE, chi - tb);
    const bool full = tile_cnt == (uint32_t)TILE;
    const uint32_t wb = tb + (uint32_t)(w * IPT * kWave + lane);
    uint32_t pk[IPT];
    uint32_t* ct = off[w];
    if (full) {
      uint32_t dg[IPT];
#pragma unroll
      for (int i = 0; i < IPT; ++i) dg[i] = digit_of<KT>(kc[i], a.shift);
      rank_items_full<IPT, false>(dg, ct, pk);
    } else {
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint32_t d = digit_of<KT>(kc[i], a.shift);
        pk[i] = (wb + i * kWave < chi) ? ((d << 16) | rank_wave_atomic<false>(d, ct)) : 0u;
      }
    }
#pragma unroll
    for (int i = 0; i < IPT; ++i) asm volatile("" : "+v"(kc[i]));
    __syncthreads();
    // next tile of the same chain and side, taken while the counts are scanned
    if (tid == 0) s_tile = msd_take(a.ctr, chain, back);
    uint32_t total = 0;
    if (tid < kRadix) {
#pragma unroll
      for (int i = 0; i < W; ++i) total += off[i][tid];
    }
    const uint32_t excl = block_excl_scan_256(tid < kRadix ? total : 0u, sums, tid);
    if (tid < kRadix) {
      uint32_t r = excl;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const uint32_t c = off[i][tid];
        off[i][tid] = r;
        r += c;
      }
      // output start of this tile's digit run, then advance the chain's
      // running offset past (front) or below (back) it
      uint32_t start;
      if (!back) {
        start = run[tid];
        run[tid] = start + total;
      } else {
        start = run[tid] - total;
        run[tid] = start;
      }
      gbase[tid] = start - excl;
    }
    __syncthreads();
    if (full) {
      uint32_t pos[IPT];
#pragma unroll
      for (int i = 0; i < IPT; ++i) pos[i] = min(off[w][pk[i] >> 16] + (pk[i] & 0xFFFFu), (uint32_t)TILE - 1);
#pragma unroll
      for (int i = 0; i < IPT; ++i) keys[pos[i]] = kc[i];
    } else {
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint32_t pos = off[w][pk[i] >> 16] + (pk[i] & 0xFFFFu);
        if (wb + i * kWave < chi && pos < (uint32_t)TILE) keys[pos] = kc[i];
      }
    }
    const uint32_t nxt = s_tile;
    if (nxt != kNoWork) load(clo, chi, nxt, kc);
    for (int i = lane; i < kRadix; i += kWave) off[w][i] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = (uint32_t)(j * kMsdBlock + tid);
      if (full || i < tile_cnt) {
        const uint32_t kk = keys[i];
        const uint32_t o = min(gbase[digit_of<KT>(kk, a.shift)] + i, a.n - 1);
        a.dst[o] = kk;
      }
    }
    __syncthreads();
    if (nxt != kNoWork) {
      tile = nxt;
      continue;
    }
    // this side of the chain is done: help another chain from its back end
    if (tid == 0) {
      uint32_t c = kNoWork, t = kNoWork;
      msd_find_back(a.ctr, a.claim, a.chains, chain + 1, c, t);
      s_chain = c;
      s_tile = t;
    }
    __syncthreads();
    if (s_tile == kNoWork) break;
    chain = s_chain;
    tile = s_tile;
    back = true;
    clo = chain_lo(chain);
    chi = chain_hi(chain);
    if (tid < kRadix) run[tid] = digit_start(chain, (uint32_t)tid, true);
    load(clo, chi, tile, kc);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// local sort of the sub-bucket batches, in place
// ---------------------------------------------------------------------------
template <int KT>
__global__ __launch_bounds__(kMsdBlock, 4) void msd_local_kernel(uint32_t* __restrict__ data,
                                                                 MsdCtl* __restrict__ ctl,
                                                                 const uint32_t* __restrict__ bucket_start,
                                                                 const uint32_t* __restrict__ sub_off,
                                                                 const uint32_t* __restrict__ leaders,
                                                                 const uint32_t* __restrict__ batch_cnt,
                                                                 const uint32_t* __restrict__ batch_base) {
  constexpr int W = kMsdBlock / kWave;
  constexpr int IPT = kLocalIpt;
  if (ctl->off) return;
  __shared__ uint32_t keys[kLocalCap];
  __shared__ uint32_t off[W][kRadix];
  __shared__ uint32_t sums[4];
  __shared__ uint32_t s_batch;
  __shared__ uint32_t s_base[kRadix + 1];

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;
  const uint32_t total = ctl->batches;
  for (int i = tid; i <= kRadix; i += kMsdBlock) s_base[i] = batch_base[i];
  for (int i = tid; i < W * kRadix; i += kMsdBlock) (&off[0][0])[i] = 0;
  const uint32_t wb = (uint32_t)(w * IPT * kWave + lane);  // wave-blocked item positions

  // bounds of batch bi: [beg, end) and whether it holds several sub-buckets
  auto bounds = [&](uint32_t bi, uint32_t& beg, uint32_t& end, bool& multi) {
    uint32_t lo = 0, hi = kRadix;  // largest b with s_base[b] <= bi
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_base[mid] <= bi) lo = mid;
      else hi = mid;
    }
    const uint32_t b = lo, k = bi - s_base[b];
    const uint32_t s0 = leaders[b * kRadix + k];
    const uint32_t s1 = k + 1 < batch_cnt[b] ? leaders[b * kRadix + k + 1] : (uint32_t)kRadix;
    beg = bucket_start[b] + sub_off[b * kRadix + s0];
    end = s1 < (uint32_t)kRadix ? bucket_start[b] + sub_off[b * kRadix + s1] : bucket_start[b + 1];
    multi = s1 - s0 > 1;
  };

  if (tid == 0) s_batch = atomicAdd(&ctl->batch_next, 1u);
  __syncthreads();
  uint32_t bi = s_batch;
  uint32_t beg = 0, end = 0;
  bool multi = false;
  uint32_t kc[IPT];
  if (bi < total) {
    bounds(bi, beg, end, multi);
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const uint32_t p = wb + i * kWave;
      kc[i] = p < end - beg ? data[beg + p] : 0u;
    }
  }
  while (bi < total) {
    const uint32_t m = end - beg;
    const int passes = multi ? 3 : 2;
    __syncthreads();
    if (tid == 0) s_batch = atomicAdd(&ctl->batch_next, 1u);
    for (int q = 0; q < passes; ++q) {
      const int shift = 8 * q;
      uint32_t k[IPT];
      if (q == 0) {
#pragma unroll
        for (int i = 0; i < IPT; ++i) k[i] = kc[i];
      } else {
#pragma unroll
        for (int i = 0; i < IPT; ++i) {
          const uint32_t p = wb + i * kWave;
          k[i] = p < m ? keys[p] : 0u;
        }
      }
      uint32_t pk[IPT];
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint32_t d = digit_of<KT>(k[i], shift);
        pk[i] = (wb + i * kWave < m) ? ((d << 16) | rank_wave_atomic<false>(d, off[w])) : 0u;
      }
      __syncthreads();
      uint32_t tot = 0;
      if (tid < kRadix) {
#pragma unroll
        for (int i = 0; i < W; ++i) tot += off[i][tid];
      }
      const uint32_t excl = block_excl_scan_256(tid < kRadix ? tot : 0u, sums, tid);
      if (tid < kRadix) {
        uint32_t r = excl;
#pragma unroll
        for (int i = 0; i < W; ++i) {
          const uint32_t c = off[i][tid];
          off[i][tid] = r;
          r += c;
        }
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint32_t pos = off[w][pk[i] >> 16] + (pk[i] & 0xFFFFu);
        if (wb + i * kWave < m && pos < kLocalCap) keys[pos] = k[i];
      }
      for (int i = lane; i < kRadix; i += kWave) off[w][i] = 0;
      if (q == 0) {
        // the batch is in LDS: load the next one into registers meanwhile
        const uint32_t nb = s_batch;
        if (nb < total) {
          uint32_t b2, e2;
          bool m2;
          bounds(nb, b2, e2, m2);
#pragma unroll
          for (int i = 0; i < IPT; ++i) {
            const uint32_t p = wb + i * kWave;
            kc[i] = p < e2 - b2 ? data[b2 + p] : 0u;
          }
        }
      }
      __syncthreads();
    }
    for (uint32_t i = (uint32_t)tid; i < m; i += kMsdBlock) data[beg + i] = keys[i];
    bi = s_batch;
    if (bi < total) bounds(bi, beg, end, multi);
  }
}

uint32_t env_u32(const char* name, uint32_t dflt) {
  const char* e = std::getenv(name);
  if (e != nullptr && e[0] != 0) {
    char* endp = nullptr;
    const unsigned long long v = std::strtoull(e, &endp, 10);
    if (endp != nullptr && *endp == 0) return (uint32_t)std::min<unsigned long long>(v, 0xFFFFFFFFull);
  }
  return dflt;
}

// Workgroups of one msd kernel resident on the device (one per CU for the
// pass kernels); cached per (kernel, device) so that graph capture issues
// no queries.
template <typename K>
uint32_t msd_resident(K kernel) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, uint32_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  const auto key = std::make_pair(reinterpret_cast<const void*>(kernel), dev);
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kMsdBlock, 0) != hipSuccess) return 0;
  const uint32_t r = (uint32_t)std::max(1, per_cu) * (uint32_t)std::max(1, cus);
  std::lock_guard<std::mutex> lock(mu);
  cache[key] = r;
  return r;
}

}  // namespace

bool msd_eligible(size_t n) {
  return n >= env_u32("MGS_MSD_MIN", MsdCfg::kMinKeys) && n <= MsdCfg::kMaxKeys &&
         env_u32("MGS_MSD", 1) != 0;
}

// Scratch of the path (at most MsdCfg::kMaxChains chains), placed from
// offset `at` of the temp buffer.
MsdLayout msd_layout(size_t at) {
  MsdLayout L{};
  size_t o = at;
  auto take = [&](size_t bytes) {
    const size_t r = o;
    o = (o + bytes + 255) / 256 * 256;
    return r;
  };
  constexpr size_t C = MsdCfg::kMaxChains;
  L.ctl = take(sizeof(MsdCtl));
  L.slices = take(C * (kMsdBins / 2) * 4);
  L.chain_hist = take(C * kRadix * 4);
  L.chain_pref = take(C * kRadix * 4);
  L.chain_bad = take(C * 4);
  L.sub_cnt = take((size_t)kMsdBins * 4);
  L.sub_off = take((size_t)kMsdBins * 4);
  L.bucket_tot = take(kRadix * 4);
  L.bucket_start = take((kRadix + 1) * 4);
  L.leaders = take((size_t)kMsdBins * 4);
  L.batch_cnt = take(kRadix * 4);
  L.batch_base = take((kRadix + 1) * 4);
  L.ctr1 = take(C * 8);
  L.ctr2 = take(kRadix * 8);
  L.claim1 = take(C * 4);
  L.claim2 = take(kRadix * 4);
  L.end = o;
  return L;
}

template <int KT>
int msd_launch(const void* kin, void* kout, void* alt, size_t n, char* t, const MsdLayout& L, uint32_t* status,
               hipStream_t s, Marks& mark) {
  const uint32_t* ki = static_cast<const uint32_t*>(kin);
  uint32_t* ko = static_cast<uint32_t*>(kout);
  uint32_t* ka = static_cast<uint32_t*>(alt);
  MsdCtl* ctl = reinterpret_cast<MsdCtl*>(t + L.ctl);
  uint32_t* slices = reinterpret_cast<uint32_t*>(t + L.slices);
  uint32_t* chain_hist = reinterpret_cast<uint32_t*>(t + L.chain_hist);
  uint32_t* chain_pref = reinterpret_cast<uint32_t*>(t + L.chain_pref);
  uint32_t* chain_bad = reinterpret_cast<uint32_t*>(t + L.chain_bad);
  uint32_t* sub_cnt = reinterpret_cast<uint32_t*>(t + L.sub_cnt);
  uint32_t* sub_off = reinterpret_cast<uint32_t*>(t + L.sub_off);
  uint32_t* bucket_tot = reinterpret_cast<uint32_t*>(t + L.bucket_tot);
  uint32_t* bucket_start = reinterpret_cast<uint32_t*>(t + L.bucket_start);
  uint32_t* leaders = reinterpret_cast<uint32_t*>(t + L.leaders);
  uint32_t* batch_cnt = reinterpret_cast<uint32_t*>(t + L.batch_cnt);
  uint32_t* batch_base = reinterpret_cast<uint32_t*>(t + L.batch_base);
  uint64_t* ctr1 = reinterpret_cast<uint64_t*>(t + L.ctr1);
  uint64_t* ctr2 = reinterpret_cast<uint64_t*>(t + L.ctr2);
  uint32_t* claim1 = reinterpret_cast<uint32_t*>(t + L.claim1);
  uint32_t* claim2 = reinterpret_cast<uint32_t*>(t + L.claim2);
  const uint32_t nn = (uint32_t)n;

  // one chain (and one workgroup of each pass) per resident pass workgroup
  uint32_t chains = msd_resident(msd_pass_kernel<KT, 1>);
  chains = std::max<uint32_t>(1, std::min<uint32_t>(chains ? chains : 256u, MsdCfg::kMaxChains));
  const size_t tiles = (n + kMsdTile - 1) / kMsdTile;
  const uint32_t chain_keys = (uint32_t)((tiles + chains - 1) / chains * kMsdTile);
  chains = (uint32_t)((n + chain_keys - 1) / chain_keys);

  mark(s, "msd_upfront");
  if ((reinterpret_cast<uintptr_t>(ki) % 16) == 0)
    msd_upfront_kernel<KT, true><<<chains, kMsdBlock, 0, s>>>(ki, nn, chain_keys, ctl, slices, chain_hist, chain_bad,
                                                              status);
  else
    msd_upfront_kernel<KT, false><<<chains, kMsdBlock, 0, s>>>(ki, nn, chain_keys, ctl, slices, chain_hist, chain_bad,
                                                               status);
  mark(s, "msd_scan");
  msd_scan_kernel<<<kRadix, kRadix, 0, s>>>(slices, chains, nn, chain_hist, chain_pref, ctl, sub_cnt, sub_off,
                                            bucket_tot, leaders, batch_cnt, chain_bad);
  mark(s, "msd_finish");
  msd_finish_kernel<<<1, kRadix, 0, s>>>(ctl, bucket_tot, bucket_start, batch_cnt, batch_base, nn, chain_keys, chains,
                                         ctr1, ctr2, claim1, claim2);
  MsdPassArgs a{};
  a.n = nn;
  a.chain_keys = chain_keys;
  a.bucket_start = bucket_start;
  a.chain_hist = chain_hist;
  a.chain_pref = chain_pref;
  a.sub_cnt = sub_cnt;
  a.sub_off = sub_off;
  a.ctl = ctl;
  mark(s, "msd_pass1");
  a.src = ki;
  a.dst = ka;
  a.chains = chains;
  a.shift = 24;
  a.ctr = ctr1;
  a.claim = claim1;
  msd_pass_kernel<KT, 1><<<chains, kMsdBlock, 0, s>>>(a);
  mark(s, "msd_pass2");
  a.src = ka;
  a.dst = ko;
  a.chains = kRadix;
  a.shift = 16;
  a.ctr = ctr2;
  a.claim = claim2;
  msd_pass_kernel<KT, 2><<<chains, kMsdBlock, 0, s>>>(a);
  mark(s, "msd_local");
  msd_local_kernel<KT><<<chains, kMsdBlock, 0, s>>>(ko, ctl, bucket_start, sub_off, leaders, batch_cnt, batch_base);
  return check_launch();
}

template int msd_launch<MGS_KEY_U32>(const void*, void*, void*, size_t, char*, const MsdLayout&, uint32_t*,
                                     hipStream_t, Marks&);
template int msd_launch<MGS_KEY_I32>(const void*, void*, void*, size_t, char*, const MsdLayout&, uint32_t*,
                                     hipStream_t, Marks&);
template int msd_launch<MGS_KEY_F32>(const void*, void*, void*, size_t, char*, const MsdLayout&, uint32_t*,
                                     hipStream_t, Marks&);

}  // namespace mgs
