// Three-pass hybrid radix sort for large arrays of 4-byte keys (u32, i32,
// f32; keys alone), the path of mgs_sort for the in-HBM sorts of 2^23 to
// 3 * 2^24 keys and of 2^29 keys and more (msd_eligible).  The plain LSD
// sort (radix_sort.hip) reads and writes the array four times; this one
// does it three times:
//
//   upfront  one read: every workgroup counts a 16-bit window of the
//            encoded keys of its range (its pass-1 chain) in a 65536-bin LDS
//            histogram of 16-bit counters, and stores it as its slice; the
//            window is the top 16 bits below the leading bits that a sample
//            of 1024 keys shares (the top 16 bits of keys that use the whole
//            range; bits 15..30 of non-negative int keys), and a key outside
//            the sample's common leading bits turns the path off
//   scan     per top byte b3: the sub-bucket sizes (b3, b2) summed over the
//            slices, their offsets inside bucket b3, the chains' prefix of
//            every b3 digit, and the batches of the local sort; a bucket or
//            sub-bucket too large for the local sort, or a 16-bit counter
//            that overflowed, turns the path off (the LSD sort runs instead)
//   finish   one workgroup: bucket starts, batch list offsets, tile counters
//   pass 1   MSD scatter by the window's high byte ("byte 3" below; byte 3
//            itself for full-range keys): workgroup c streams chain c (a fixed
//            range of the input) tile by tile with running per-digit output
//            offsets -- the chain's digit bases come from the upfront
//            histograms, so there is no look-back between workgroups
//   pass 2   scatter by the window's low byte ("byte 2") inside each byte-3
//            bucket: workgroup b streams
//            bucket b the same way, bases from the sub-bucket offsets
//   local    every batch of consecutive sub-buckets (at most kLocalCap keys,
//            one window value, or several of one bucket) is read once,
//            sorted by its low 16 bits (one sub-bucket: its keys agree on
//            every bit above them) or 24 bits (one bucket) in LDS and written
//            back in place
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
// kOffSample: a bucket holds more than kSampleCrowd of the 1024 sampled keys
// (5 times its share: far beyond the scan's limit of twice the share, which
// uniform keys reach in the sample with a chance of about 1e-6), so the
// upfront read is skipped and the scan and finish kernels return at once.
enum : uint32_t { kOffOverflow = 1u, kOffSubBucket = 2u, kOffBucket = 4u, kOffSample = 8u };
constexpr uint32_t kSampleCrowd = 20;


// ---------------------------------------------------------------------------
// upfront
// ---------------------------------------------------------------------------
template <int KT, bool VEC>
__global__ __launch_bounds__(kMsdBlock) void msd_upfront_kernel(const uint32_t* __restrict__ keys, uint32_t n,
                                                               uint32_t chain_keys, MsdCtl* __restrict__ ctl,
                                                               uint32_t* __restrict__ slices,
                                                               uint32_t* __restrict__ chain_hist,
                                                               uint32_t* __restrict__ chain_bad,
                                                               uint32_t* __restrict__ xcd_of,
                                                               uint32_t* __restrict__ desc, uint32_t desc_tiles,
                                                               uint32_t* __restrict__ status) {
  // 65536 16-bit counters, two per word: bin b in word b >> 1, half b & 1
  __shared__ uint32_t h[kMsdBins / 2];
  __shared__ uint32_t s_rows[kRadix];
  __shared__ uint32_t s_runs, s_stray;
  __shared__ uint32_t s_and[kMsdBlock / kWave], s_or[kMsdBlock / kWave];
  __shared__ uint32_t s_bh[kRadix];
  __shared__ uint32_t s_crowd;
  const int tid = threadIdx.x;
  const uint32_t c = blockIdx.x;
  const int lane = tid & (kWave - 1), w = tid / kWave;
  for (int i = tid; i < kMsdBins / 2; i += kMsdBlock) h[i] = 0;
  if (tid < kRadix) s_bh[tid] = 0;
  if (tid == 0) {
    xcd_of[c] = xcc_id();
    s_stray = 0;
    s_crowd = 0;
  }
  // the window: every workgroup reads the same sample of 1024 keys, evenly
  // spaced over the input (n >= 2), and derives the same bits from it
  const uint32_t e_sample = Codec<KT>::enc(keys[((uint64_t)tid * (n - 1)) / (kMsdBlock - 1)]);
  {
    const uint32_t e = e_sample;
    uint32_t ea = e, eo = e;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ea &= __shfl_xor(ea, o, kWave);
      eo |= __shfl_xor(eo, o, kWave);
    }
    if (lane == 0) {
      s_and[w] = ea;
      s_or[w] = eo;
    }
  }
  __syncthreads();
  uint32_t s_a = ~0u, s_o = 0u;
#pragma unroll
  for (int i = 0; i < kMsdBlock / kWave; ++i) {
    s_a &= s_and[i];
    s_o |= s_or[i];
  }
  const uint32_t common = (s_a ^ s_o) ? (uint32_t)__clz(s_a ^ s_o) : 32u;  // leading bits the sample shares
  const uint32_t low = 16u - min(common, 16u);
  const uint32_t hi_mask = low < 16u ? ~0u << (low + 16u) : 0u;  // the shared bits above the window
  const uint32_t hi_val = s_a & hi_mask;
  // the sample's keys per bucket (the window's high byte): a crowded bucket
  // hands the sort to the LSD passes before the read
  atomicAdd(&s_bh[(e_sample >> (low + 8u)) & 0xFFu], 1u);
  __syncthreads();
  if (tid < kRadix && s_bh[tid] > kSampleCrowd) s_crowd = 1u;
  __syncthreads();
  if (s_crowd) {
    if (c == 0 && tid == 0) ctl->off = kOffSample;  // (the LSD sort clears the status word)
    return;
  }
  if (c == 0 && tid == 0) {  // nothing else writes these during this kernel
    *status = 0;
    ctl->off = 0;
    ctl->xcd_mask = 0;
    ctl->low = low;
  }
  auto window = [&](uint32_t enc) { return (enc >> low) & 0xFFFFu; };
  uint32_t stray = 0;  // bits above the window where a key differs from the sample
  {
    // look-back descriptors of both passes: the tiles of the input read as
    // "empty" by pass 1 (status rotation 0) and, once pass 1 has left them
    // inclusive, by pass 2 (rotation 1); the up to 256 further tiles of
    // pass 2 (one partial tile per bucket) start as "empty" under rotation 1
    uint4* d4 = reinterpret_cast<uint4*>(desc);
    const size_t words = (size_t)desc_tiles * kRadix;
    const size_t total = words + (size_t)(kRadix + kXcds) * kRadix;
    const size_t g = (size_t)c * kMsdBlock + tid, st = (size_t)gridDim.x * kMsdBlock;
    const uint32_t empty1 = Desc<uint32_t>::make(kStEmpty, 0u, 1);
    for (size_t i = g; i < total / 4; i += st) {
      const uint32_t v = 4 * i < words ? 0u : empty1;
      d4[i] = make_uint4(v, v, v, v);
    }
  }
  __syncthreads();
  const size_t lo = (size_t)c * chain_keys;
  const size_t hi = std::min<size_t>((size_t)n, lo + chain_keys);
  auto add = [&](uint32_t b, uint32_t k) {
    __hip_atomic_fetch_add(&h[b >> 1], k << ((b & 1u) << 4), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  auto count = [&](uint32_t k) {
    const uint32_t e = Codec<KT>::enc(k);
    stray |= (e & hi_mask) ^ hi_val;
    add(window(e), 1u);
  };
  // four consecutive keys per lane: one add when they share a bin, and one
  // add for all the lanes whose four keys fall in the first lane's bin (runs
  // of sorted keys, which would otherwise put every lane of the wavefront on
  // one LDS counter); such keys are counted in `run_keys`
  uint32_t run_keys = 0;
  auto count4 = [&](const uint4& v) {
    const uint32_t e0 = Codec<KT>::enc(v.x), e1 = Codec<KT>::enc(v.y);
    const uint32_t e2 = Codec<KT>::enc(v.z), e3 = Codec<KT>::enc(v.w);
    stray |= ((e0 & hi_mask) ^ hi_val) | ((e1 & hi_mask) ^ hi_val) | ((e2 & hi_mask) ^ hi_val) |
             ((e3 & hi_mask) ^ hi_val);
    const uint32_t b0 = window(e0), b1 = window(e1);
    const uint32_t b2 = window(e2), b3 = window(e3);
    const bool same = (b0 == b1) & (b0 == b2) & (b0 == b3);
    const bool in = same && b0 == readfirstlane(b0);
    const uint64_t grp = ballot64(in);
    const uint32_t k = 4u * (uint32_t)__popcll(grp);
    if (mbcnt64(ballot64(true)) == 0) run_keys += k;
    if (in) {
      if (mbcnt64(grp) == 0) add(b0, k);
    } else if (same) {
      add(b0, 4u);
    } else {
      add(b0, 1u);
      add(b1, 1u);
      add(b2, 1u);
      add(b3, 1u);
    }
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
        for (int u = 0; u < U; ++u) count4(v[u]);
      }
      for (; i < nv; i += kMsdBlock) count4(kv[i]);
      for (size_t k = lo + nv * 4 + tid; k < hi; k += kMsdBlock) count(keys[k]);
    } else {
      for (size_t k = lo + tid; k < hi; k += kMsdBlock) count(keys[k]);
    }
  }
  if (tid == 0) s_runs = 0;
  __syncthreads();
  if (run_keys) atomicAdd(&s_runs, run_keys);
  if (stray) s_stray = 1u;
  // byte-3 histogram of the chain: row r is the 128 words [128 r, 128 r + 128)
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
    // range, and the path is turned off; so it is by a key outside the
    // sample's shared leading bits
    uint32_t s = s_rows[tid] + s_rows[tid + 64] + s_rows[tid + 128] + s_rows[tid + 192];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    if (tid == 0)
      chain_bad[c] = (s != (uint32_t)(hi > lo ? hi - lo : 0) || s_stray ? 1u : 0u) | (s_runs << 1);
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
                                                       const uint32_t* __restrict__ chain_bad,
                                                       const uint32_t* __restrict__ xcd_of) {
  if (ctl->off & kOffSample) return;
  __shared__ uint32_t sums[4];
  __shared__ uint32_t s_carry;
  const uint32_t b = blockIdx.x, d = threadIdx.x;
  const uint32_t s = b * kRadix + d;
  if (b == 0) {
    uint32_t m = 0, runs = 0;
    for (uint32_t c = d; c < chains; c += kRadix) {
      const uint32_t cb = chain_bad[c];
      if (cb & 1u) atomicOr(&ctl->off, kOffOverflow);
      runs += cb >> 1;
      m |= 1u << (xcd_of[c] & 31u);
    }
    atomicOr(&ctl->xcd_mask, m);
    const uint32_t before = block_excl_scan_256(runs, sums, (int)d);
    if (d == kRadix - 1) ctl->runs = before + runs > n / 16 ? 1u : 0u;
    __syncthreads();  // every thread has read the scan's partial sums
  }
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
                                                         uint32_t* __restrict__ batch_base,
                                                         uint32_t chains1, uint32_t tpc1,
                                                         uint32_t* __restrict__ tile_base1,
                                                         uint32_t* __restrict__ tile_base2,
                                                         uint32_t* __restrict__ counters) {
  if (ctl->off & kOffSample) return;
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
  __syncthreads();
  // descriptor tiles: pass 1 chain C from C * tpc1; pass 2 bucket b after
  // the tiles of the buckets before it
  const uint32_t nt = (bt + kMsdTile - 1) / kMsdTile;
  tile_base2[d] = block_excl_scan_256(nt, sums, (int)d);
  if (d < chains1) tile_base1[d] = d * tpc1;
  counters[d] = 0;  // pass 2
  if (d < kXcds) counters[kRadix + d] = 0;  // pass 1
}

// ---------------------------------------------------------------------------
// passes 1 and 2: onesweep passes over chains of tiles with decoupled
// look-back inside each chain (the look-back of the LSD pass kernel).
//   pass 1  8 chains: equal ranges of the input; the digit bases of chain C
//           are the byte-3 bucket starts plus the keys of that digit in
//           the ranges before it (the upfront's per-range histograms)
//   pass 2  256 chains: the byte-3 buckets; digit bases = sub-bucket starts
// Chain c is served by the workgroups of XCD c mod 8 (when the upfront ran
// on all eight XCDs), which start together on their XCD's first chain and
// move on together, so at any time the 32 workgroups of an XCD work on
// consecutive tiles of one chain: every digit run of the chain's output is
// then written by consecutive tiles at about the same time.
// ---------------------------------------------------------------------------
struct MsdPassArgs {
  const uint32_t* src;
  uint32_t* dst;
  uint32_t n;
  uint32_t chains;             // 8 (pass 1) or 256 (pass 2)
  const uint32_t* bucket_start;
  const uint32_t* chain_pref;  // pass 1: prefix over the upfront's ranges per digit
  uint32_t ranges_per_chain;   // pass 1: upfront ranges per chain
  uint32_t chain_keys;         // pass 1: keys per chain
  const uint32_t* sub_off;     // pass 2: sub-bucket offsets inside each bucket
  const uint32_t* tile_base;   // first descriptor tile of each chain
  uint32_t* counters;          // tiles taken per chain
  uint32_t* desc;              // look-back descriptors, 256 per tile
  int parity;                  // descriptor status rotation of this pass
  const MsdCtl* ctl;
  uint32_t* status;
  uint32_t spin_limit;
};

__device__ __forceinline__ void msd_publish(uint32_t* p, uint32_t v, bool xcd_local) {
  if (xcd_local)
    st_wg(p, v);
  else
    st_agent(p, v);
}

template <int KT, int PASS>
__global__ __launch_bounds__(kMsdBlock, 4) void msd_pass_kernel(MsdPassArgs a) {
  constexpr int W = kMsdBlock / kWave;
  constexpr int IPT = kMsdIpt;
  constexpr int TILE = kMsdTile;
  using D = uint32_t;
  if (a.ctl->off) return;
  const bool runs = a.ctl->runs != 0;
  const int shift = (int)a.ctl->low + (PASS == 1 ? 8 : 0);  // the window's high or low byte

  __shared__ uint32_t keys[TILE];
  __shared__ uint32_t off[W][kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint32_t s_prefix[kRadix];
  __shared__ uint32_t sums[4];
  __shared__ uint32_t s_tile;

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;
  const bool lb_wave = tid >= kRadix && tid < 2 * kRadix;
  const int taker = kRadix;
  for (int i = tid; i < W * kRadix; i += kMsdBlock) (&off[0][0])[i] = 0;

  auto chain_lo = [&](uint32_t c) -> uint32_t {
    if constexpr (PASS == 1) return c * a.chain_keys;
    else return a.bucket_start[c];
  };
  auto chain_hi = [&](uint32_t c) -> uint32_t {
    if constexpr (PASS == 1) return (uint32_t)std::min<size_t>((size_t)(c + 1) * a.chain_keys, a.n);
    else return a.bucket_start[c + 1];
  };
  // thread d < 256: output position of the chain's first key of digit d
  auto digit_base = [&](uint32_t c, uint32_t d) -> uint32_t {
    if constexpr (PASS == 1) return a.bucket_start[d] + a.chain_pref[(size_t)c * a.ranges_per_chain * kRadix + d];
    else return a.bucket_start[c] + a.sub_off[c * kRadix + d];
  };

  const bool xcd_local = a.chains % kXcds == 0 && a.ctl->xcd_mask == (1u << kXcds) - 1;
  uint32_t my_chain = blockIdx.x % a.chains, chains_left = a.chains, chain_step = 1;
  if (xcd_local) {
    my_chain = xcc_id();
    chains_left = a.chains / kXcds;
    chain_step = kXcds;
  }
  // next tile: (chain, tile inside the chain) packed as chain << 20 | tile
  auto take = [&]() -> uint32_t {
    while (chains_left > 0) {
      const uint32_t t = atomicAdd(&a.counters[my_chain], 1u);
      const uint32_t lo = chain_lo(my_chain), hi = chain_hi(my_chain);
      if (t < (hi - lo + TILE - 1) / TILE) return (my_chain << 20) | t;
      my_chain += chain_step;
      if (my_chain >= a.chains) my_chain -= a.chains;
      --chains_left;
    }
    return kNoWork;
  };

  if (tid == taker) s_tile = take();
  __syncthreads();
  uint32_t cur = s_tile;
  uint32_t cur_chain = 0xFFFFFFFFu, cur_base = 0, clo = 0, chi = 0;
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
  if (cur != kNoWork) load(chain_lo(cur >> 20), chain_hi(cur >> 20), cur & 0xFFFFFu, kc);

  while (cur != kNoWork) {
    const uint32_t chain = cur >> 20, lt = cur & 0xFFFFFu;
    if (chain != cur_chain) {
      clo = chain_lo(chain);
      chi = chain_hi(chain);
    }
    const uint32_t tb = clo + lt * (uint32_t)TILE;
    const uint32_t tile_cnt = min((uint32_t)TILE, chi - tb);
    const bool full = tile_cnt == (uint32_t)TILE;
    const uint32_t wb = tb + (uint32_t)(w * IPT * kWave + lane);
    uint32_t pk[IPT];
    uint32_t* ct = off[w];
    if (full) {
      uint32_t dg[IPT];
#pragma unroll
      for (int i = 0; i < IPT; ++i) dg[i] = digit_of<KT>(kc[i], shift);
      if (runs)
        rank_items_full<IPT, true>(dg, ct, pk);
      else
        rank_items_full<IPT, false>(dg, ct, pk);
    } else {
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint32_t d = digit_of<KT>(kc[i], shift);
        pk[i] = (wb + i * kWave < chi) ? ((d << 16) | rank_wave_atomic<false>(d, ct)) : 0u;
      }
    }
#pragma unroll
    for (int i = 0; i < IPT; ++i) asm volatile("" : "+v"(kc[i]));
    __syncthreads();

    uint32_t total = 0;
    D* chain_desc = a.desc + (size_t)a.tile_base[chain] * kRadix;
    D win[kPassWindow];
    uint32_t taken = kNoWork;
    if (lb_wave) {
      const uint32_t d = (uint32_t)tid - kRadix;
      if (tid == taker) taken = take();
#pragma unroll
      for (int i = 0; i < W; ++i) total += off[i][d];
      msd_publish(chain_desc + (size_t)lt * kRadix + d,
                  Desc<D>::make(lt == 0 ? kStInclusive : kStAggregate, (D)total, a.parity), xcd_local);
      if (lt > 0) lookback_load(chain_desc, (int64_t)lt - 1, d, win);
    } else if (tid < kRadix) {
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
      if (chain != cur_chain) cur_base = digit_base(chain, (uint32_t)tid);
      s_prefix[tid] = cur_base - excl;
    }
    cur_chain = chain;
    if (tid == taker) s_tile = taken;
    __syncthreads();
    int64_t lb_t = (int64_t)lt - 1;
    uint32_t lb_prefix = 0;
    bool lb_done = false;
    if (lb_wave && lt > 0)
      lb_done = lookback_step(chain_desc, lb_t, (uint32_t)tid - kRadix, a.parity, win, lb_prefix);
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
    if (nxt != kNoWork) {
      const uint32_t nc = nxt >> 20;
      load(chain_lo(nc), chain_hi(nc), nxt & 0xFFFFFu, kc);
    }
    for (int i = lane; i < kRadix; i += kWave) off[w][i] = 0;
    if (lb_wave) {
      const uint32_t d = (uint32_t)tid - kRadix;
      if (lt > 0) {
        bool gave_up = false;
        const uint32_t prefix =
            lb_done ? lb_prefix : lookback_from(chain_desc, lb_t, d, a.parity, win, lb_prefix, a.spin_limit, gave_up);
        if (gave_up) atomicOr(a.status, kStatusLookback);
        msd_publish(chain_desc + (size_t)lt * kRadix + d, Desc<D>::make(kStInclusive, (D)(prefix + total), a.parity),
                    xcd_local);
        gbase[d] = s_prefix[d] + prefix;
      } else {
        gbase[d] = s_prefix[d];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = (uint32_t)(j * kMsdBlock + tid);
      if (full || i < tile_cnt) {
        const uint32_t kk = keys[i];
        const uint32_t o = min(gbase[digit_of<KT>(kk, shift)] + i, a.n - 1);
        a.dst[o] = kk;
      }
    }
    __syncthreads();
    cur = nxt;
  }
}

// ---------------------------------------------------------------------------
// batch table: workgroup b, thread k < batch_cnt[b] writes batch
// batch_base[b] + k of bucket b as {first key, key count | several
// sub-buckets << 31}
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void msd_batches_kernel(const MsdCtl* __restrict__ ctl,
                                                          const uint32_t* __restrict__ bucket_start,
                                                          const uint32_t* __restrict__ sub_off,
                                                          const uint32_t* __restrict__ leaders,
                                                          const uint32_t* __restrict__ batch_cnt,
                                                          const uint32_t* __restrict__ batch_base,
                                                          uint2* __restrict__ batches) {
  if (ctl->off) return;
  const uint32_t b = blockIdx.x, k = threadIdx.x;
  const uint32_t cnt = batch_cnt[b];
  if (k >= cnt) return;
  const uint32_t s0 = leaders[b * kRadix + k];
  const uint32_t s1 = k + 1 < cnt ? leaders[b * kRadix + k + 1] : (uint32_t)kRadix;
  const uint32_t beg = bucket_start[b] + sub_off[b * kRadix + s0];
  const uint32_t end = s1 < (uint32_t)kRadix ? bucket_start[b] + sub_off[b * kRadix + s1] : bucket_start[b + 1];
  batches[batch_base[b] + k] = make_uint2(beg, (end - beg) | (s1 - s0 > 1 ? 0x80000000u : 0u));
}

// ---------------------------------------------------------------------------
// local sort of the batches, in place: workgroup g sorts batches g, g + G,
// g + 2G, ... (G workgroups).  A batch of one sub-bucket is sorted by bytes
// 0 and 1 (its top 16 bits are one value), a batch of several by bytes 0, 1
// and 2 (one bucket: byte 3 is one value).  The next batch's keys are loaded
// into registers while the current one is sorted in LDS.  Block 0 also
// checks that both passes took every tile of every chain.
// ---------------------------------------------------------------------------
template <int KT>
__global__ __launch_bounds__(kMsdBlock, 4) void msd_local_kernel(uint32_t* __restrict__ data,
                                                                 const MsdCtl* __restrict__ ctl,
                                                                 const uint2* __restrict__ batches,
                                                                 const uint32_t* __restrict__ counters,
                                                                 const uint32_t* __restrict__ bucket_start,
                                                                 uint32_t chains1, uint32_t chain_keys1, uint32_t n,
                                                                 uint32_t* __restrict__ status) {
  constexpr int W = kMsdBlock / kWave;
  constexpr int IPT = kLocalIpt;
  if (ctl->off) return;
  // keys[kLocalCap + lane]: where the lanes without a key (beyond the
  // batch) store, so that every LDS access is unconditional
  __shared__ uint32_t keys[kLocalCap + kWave];
  // off[w][kRadix + lane]: the counters lanes without a key count into (one
  // per lane: a shared one would serialise the atomics)
  __shared__ uint32_t off[W][kRadix + kWave];
  __shared__ uint32_t sums[4];

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;
  if (blockIdx.x == 0) {
    // counters: [0, 256) pass 2 per bucket, [256, 264) pass 1 per chain
    if (tid < kRadix) {
      const uint32_t len = bucket_start[tid + 1] - bucket_start[tid];
      if (counters[tid] < (len + kMsdTile - 1) / kMsdTile) atomicOr(status, kStatusChainUnserved);
    } else if (tid < kRadix + (int)chains1) {
      const uint32_t c = tid - kRadix;
      const uint32_t lo = c * chain_keys1;
      const uint32_t len = min(n - lo, chain_keys1);
      if (counters[tid] < (len + kMsdTile - 1) / kMsdTile) atomicOr(status, kStatusChainUnserved);
    }
  }
  const uint32_t total = ctl->batches;
  const uint32_t G = gridDim.x;
  for (int i = tid; i < W * (kRadix + kWave); i += kMsdBlock) (&off[0][0])[i] = 0;
  // item i of a lane sits at wb + 64 i: wavefront w holds the positions
  // [w IPT 64, (w+1) IPT 64), in the order the ranking keeps stable (a
  // batch smaller than kLocalCap leaves the last wavefronts idle: measured
  // faster than spreading it over all of them, which gives each wavefront
  // fewer LDS atomics in flight)
  const uint32_t wb = (uint32_t)(w * IPT * kWave + lane);
  // (unconditional loads: positions beyond the batch read its last key, or
  // key 0 for an empty batch, and are ignored)
  auto load_batch = [&](uint2 bt, uint32_t (&k)[IPT]) {
    const uint32_t m = bt.y & 0x7FFFFFFFu;
    const uint32_t* src = data + (m ? bt.x : 0u);
    const uint32_t last = m ? m - 1 : 0u;
#pragma unroll
    for (int i = 0; i < IPT; ++i) k[i] = src[min(wb + i * kWave, last)];
  };

  uint32_t bi = blockIdx.x;
  uint2 cur = bi < total ? batches[bi] : make_uint2(0, 0);
  uint32_t kc[IPT];
  load_batch(cur, kc);

  // one LSD pass over byte shift/8 of the m keys in k into LDS, in order;
  // branch-free, so that each wavefront's LDS operations issue back to back
  auto lds_pass = [&](const uint32_t (&k)[IPT], int shift, uint32_t m) {
    // a wavefront whose positions all lie beyond the batch skips its LDS
    // work (a wave-uniform branch)
    const bool wave_has_keys = (uint32_t)(w * IPT * kWave) < m;
    uint32_t dg[IPT], pk[IPT];
#pragma unroll
    for (int i = 0; i < IPT; ++i)
      dg[i] = wb + i * kWave < m ? digit_of<KT>(k[i], shift) : (uint32_t)(kRadix + lane);
    if (wave_has_keys) rank_items_full<IPT, false>(dg, off[w], pk);
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
    if (wave_has_keys) {
      uint32_t pos[IPT];
#pragma unroll
      for (int i = 0; i < IPT; ++i) pos[i] = off[w][pk[i] >> 16];  // lanes without a key read their spare counter
#pragma unroll
      for (int i = 0; i < IPT; ++i)
        pos[i] = (pk[i] >> 16) < (uint32_t)kRadix ? min(pos[i] + (pk[i] & 0xFFFFu), kLocalCap - 1)
                                                  : kLocalCap + (uint32_t)lane;
#pragma unroll
      for (int i = 0; i < IPT; ++i) keys[pos[i]] = k[i];
    }
    // every row is cleared: the scan wrote running offsets into the rows of
    // the wavefronts without keys as well
    for (int i = lane; i < kRadix + kWave; i += kWave) off[w][i] = 0;
  };

  while (bi < total) {
    const uint32_t beg = cur.x, m = cur.y & 0x7FFFFFFFu;
    const int passes = (cur.y >> 31) ? 3 : 2;
    const uint32_t nb = bi + G;
    const uint2 nxt = nb < total ? batches[nb] : make_uint2(0, 0);
    __syncthreads();
    lds_pass(kc, 0, m);
    // the batch is in LDS: load the next one into registers meanwhile
    if (nb < total) load_batch(nxt, kc);
    __syncthreads();
    for (int q = 1; q < passes; ++q) {
      uint32_t k[IPT];
#pragma unroll
      for (int i = 0; i < IPT; ++i) k[i] = keys[min(wb + i * kWave, kLocalCap + kWave - 1)];
      lds_pass(k, 8 * q, m);
      __syncthreads();
    }
    // write-out: all LDS reads first, then the stores
    {
      uint32_t v[IPT];
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t i = (uint32_t)(j * kMsdBlock + tid);
        v[j] = i < m ? keys[i] : 0u;
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t i = (uint32_t)(j * kMsdBlock + tid);
        if (i < m) data[beg + i] = v[j];
      }
    }
    bi = nb;
    cur = nxt;
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
  if (env_u32("MGS_MSD", 1) == 0 || n > MsdCfg::kMaxKeys) return false;
  const char* e = std::getenv("MGS_MSD_MIN");
  if (e != nullptr && e[0] != 0) return n >= env_u32("MGS_MSD_MIN", 0);
  return (n >= MsdCfg::kMinKeys && n < MsdCfg::kGapLo) || n >= MsdCfg::kGapHi;
}

// Scratch of the path (at most MsdCfg::kMaxChains upfront ranges) from
// offset `at` of temp: the look-back descriptors of its passes first (they
// may overlap the LSD sort's, which only a sort the path turned itself off
// for uses), then its tables; the control block goes at `ctl_at`, which must
// lie outside what the LSD sort clears or writes.
MsdLayout msd_layout(size_t n, size_t at, size_t ctl_at) {
  MsdLayout L{};
  size_t o = at;
  auto take = [&](size_t bytes) {
    const size_t r = o;
    o = (o + bytes + 255) / 256 * 256;
    return r;
  };
  constexpr size_t C = MsdCfg::kMaxChains;
  const size_t tiles = (n + kMsdTile - 1) / kMsdTile;
  L.ctl = ctl_at;
  L.desc = take((tiles + kRadix + kXcds) * kRadix * 4);
  L.slices = take(C * (kMsdBins / 2) * 4);
  L.chain_hist = take(C * kRadix * 4);
  L.chain_pref = take(C * kRadix * 4);
  L.chain_bad = take(C * 4);
  L.xcd_of = take(C * 4);
  L.sub_cnt = take((size_t)kMsdBins * 4);
  L.sub_off = take((size_t)kMsdBins * 4);
  L.bucket_tot = take(kRadix * 4);
  L.bucket_start = take((kRadix + 1) * 4);
  L.leaders = take((size_t)kMsdBins * 4);
  L.batch_cnt = take(kRadix * 4);
  L.batch_base = take((kRadix + 1) * 4);
  L.batches = take((size_t)kMsdBins * 8);
  L.tile_base1 = take(kXcds * 4);
  L.tile_base2 = take(kRadix * 4);
  L.counters = take((kRadix + kXcds) * 4);
  L.end = o;
  return L;
}

template <int KT>
int msd_launch(const void* kin, void* kout, void* alt, size_t n, char* t, const MsdLayout& L, uint32_t* status,
               uint32_t spin_limit, hipStream_t s, Marks& mark) {
  const uint32_t* ki = static_cast<const uint32_t*>(kin);
  uint32_t* ko = static_cast<uint32_t*>(kout);
  uint32_t* ka = static_cast<uint32_t*>(alt);
  auto at = [&](size_t o) { return reinterpret_cast<uint32_t*>(t + o); };
  MsdCtl* ctl = reinterpret_cast<MsdCtl*>(t + L.ctl);
  uint32_t* desc = at(L.desc);
  uint32_t* slices = at(L.slices);
  uint32_t* chain_hist = at(L.chain_hist);
  uint32_t* chain_pref = at(L.chain_pref);
  uint32_t* chain_bad = at(L.chain_bad);
  uint32_t* xcd_of = at(L.xcd_of);
  uint32_t* sub_cnt = at(L.sub_cnt);
  uint32_t* sub_off = at(L.sub_off);
  uint32_t* bucket_tot = at(L.bucket_tot);
  uint32_t* bucket_start = at(L.bucket_start);
  uint32_t* leaders = at(L.leaders);
  uint32_t* batch_cnt = at(L.batch_cnt);
  uint32_t* batch_base = at(L.batch_base);
  uint2* batches = reinterpret_cast<uint2*>(t + L.batches);
  uint32_t* tile_base1 = at(L.tile_base1);
  uint32_t* tile_base2 = at(L.tile_base2);
  uint32_t* counters = at(L.counters);
  const uint32_t nn = (uint32_t)n;

  // upfront ranges: up to kMaxChains, a whole number of tiles each, grouped
  // 32 to a pass-1 chain (8 chains at full size)
  const uint32_t tiles = (uint32_t)((n + kMsdTile - 1) / kMsdTile);
  const uint32_t tpr = (tiles + MsdCfg::kMaxChains - 1) / MsdCfg::kMaxChains;
  const uint32_t range_keys = tpr * kMsdTile;
  const uint32_t ranges = (uint32_t)((n + range_keys - 1) / range_keys);
  constexpr uint32_t kRangesPerChain = MsdCfg::kMaxChains / kXcds;
  const uint32_t chains1 = (ranges + kRangesPerChain - 1) / kRangesPerChain;
  // workgroups of the passes and the local sort: as many as are resident
  uint32_t grid = msd_resident(msd_pass_kernel<KT, 1>);
  grid = grid ? grid : 256u;

  mark(s, "msd_upfront");
  if ((reinterpret_cast<uintptr_t>(ki) % 16) == 0)
    msd_upfront_kernel<KT, true><<<ranges, kMsdBlock, 0, s>>>(ki, nn, range_keys, ctl, slices, chain_hist, chain_bad,
                                                              xcd_of, desc, tiles, status);
  else
    msd_upfront_kernel<KT, false><<<ranges, kMsdBlock, 0, s>>>(ki, nn, range_keys, ctl, slices, chain_hist, chain_bad,
                                                               xcd_of, desc, tiles, status);
  mark(s, "msd_scan");
  msd_scan_kernel<<<kRadix, kRadix, 0, s>>>(slices, ranges, nn, chain_hist, chain_pref, ctl, sub_cnt, sub_off,
                                            bucket_tot, leaders, batch_cnt, chain_bad, xcd_of);
  mark(s, "msd_finish");
  msd_finish_kernel<<<1, kRadix, 0, s>>>(ctl, bucket_tot, bucket_start, batch_cnt, batch_base, chains1,
                                         kRangesPerChain * tpr, tile_base1, tile_base2, counters);
  MsdPassArgs a{};
  a.n = nn;
  a.bucket_start = bucket_start;
  a.chain_pref = chain_pref;
  a.ranges_per_chain = kRangesPerChain;
  a.chain_keys = kRangesPerChain * range_keys;
  a.sub_off = sub_off;
  a.desc = desc;
  a.ctl = ctl;
  a.status = status;
  a.spin_limit = spin_limit;
  mark(s, "msd_pass1");
  a.src = ki;
  a.dst = ka;
  a.chains = chains1;
  a.tile_base = tile_base1;
  a.counters = counters + kRadix;
  a.parity = 0;
  msd_pass_kernel<KT, 1><<<std::min(grid, tiles), kMsdBlock, 0, s>>>(a);
  mark(s, "msd_pass2");
  a.src = ka;
  a.dst = ko;
  a.chains = kRadix;
  a.tile_base = tile_base2;
  a.counters = counters;
  a.parity = 1;
  msd_pass_kernel<KT, 2><<<grid, kMsdBlock, 0, s>>>(a);
  mark(s, "msd_batches");
  msd_batches_kernel<<<kRadix, kRadix, 0, s>>>(ctl, bucket_start, sub_off, leaders, batch_cnt, batch_base, batches);
  mark(s, "msd_local");
  msd_local_kernel<KT><<<grid, kMsdBlock, 0, s>>>(ko, ctl, batches, counters, bucket_start, chains1,
                                                  kRangesPerChain * range_keys, nn, status);
  return check_launch();
}

template int msd_launch<MGS_KEY_U32>(const void*, void*, void*, size_t, char*, const MsdLayout&, uint32_t*, uint32_t,
                                     hipStream_t, Marks&);
template int msd_launch<MGS_KEY_I32>(const void*, void*, void*, size_t, char*, const MsdLayout&, uint32_t*, uint32_t,
                                     hipStream_t, Marks&);
template int msd_launch<MGS_KEY_F32>(const void*, void*, void*, size_t, char*, const MsdLayout&, uint32_t*, uint32_t,
                                     hipStream_t, Marks&);

}  // namespace mgs
