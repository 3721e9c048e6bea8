// Onesweep LSD radix sort kernels and their launch sequence (see
// radix_sort.hpp for the algorithm outline).
#include <cstdlib>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include "internal.hpp"
#include "msd_sort.hpp"
#include "radix_device.hpp"

namespace mgs {


// ---------------------------------------------------------------------------
// Pipeline (one call, fixed launch sequence, no host synchronisation):
//
//   init      zero histograms and tile counters, seed the OR/AND reductions
//   upfront   one read: per-chain histograms of byte 0 of the encoded keys,
//             bitwise OR and AND of all encoded keys (byte b is trivial -- one
//             value for every key -- iff OR and AND agree on it); clears the
//             look-back descriptors
//   pass p    p = 0..P-1, persistent workgroups; a trivial byte returns at
//             once.  The pass's tiles form kChains chains of consecutive
//             tiles, each served by the workgroups of one XCD when the
//             device shows all eight (XCD-local look-back, below), else
//             workgroup b starts on chain b mod chains and then helps the
//             others.  Per tile: rank the keys per wavefront (one LDS atomic
//             each); waves 4..7 publish the tile's digit counts and start the
//             look-back over the chain's predecessors while waves 0..3 scan
//             the counts; every wavefront scatters its keys into the LDS
//             tile and loads its part of the next tile; waves 4..7 finish
//             the look-back; then each digit run is written with consecutive
//             lanes on consecutive addresses.  On the way out every key is
//             counted in the histogram of the next non-trivial byte, per
//             chain of its output position -- exactly the per-chain
//             histograms the next pass starts from.
//   fixup     copies the result into keys_out when the pass count left it in
//             the scratch buffer
// ---------------------------------------------------------------------------

constexpr uint32_t kNoTile = 0xFFFFFFFFu;

// gate: NULL, or the decision word of the three-pass path launched before
// this sort (msd_sort.hip); the LSD kernels then run only when it is nonzero
// (the path turned itself off).
__device__ __forceinline__ bool gated_off(const uint32_t* gate) { return gate != nullptr && *gate == 0; }

__global__ void init_kernel(uint32_t* __restrict__ hist, uint32_t words, uint32_t* __restrict__ counters,
                            uint32_t counter_words, uint64_t* __restrict__ red, uint32_t* __restrict__ status,
                            const uint32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t i = g; i < words; i += gridDim.x * blockDim.x) hist[i] = 0;
  for (uint32_t i = g; i < counter_words; i += gridDim.x * blockDim.x) counters[i] = 0;
  if (g == 0) {
    *status = 0;
    red[0] = 0;     // OR identity
    red[1] = ~0ull;  // AND identity
    red[2] = 0;      // XCDs seen by the upfront kernel
    red[3] = 0;      // crowded bytes of the upfront kernel's sample
  }
}

template <typename Bits>
__device__ __forceinline__ Bits wave_or(Bits v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, kWave);
  return v;
}
template <typename Bits>
__device__ __forceinline__ Bits wave_and(Bits v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v &= __shfl_xor(v, o, kWave);
  return v;
}

// Chain of workgroup b in a chain-major split of `grid` workgroups, and the
// workgroup's rank among those serving that chain.
__device__ __forceinline__ void block_chain(uint32_t b, uint32_t grid, uint32_t chains, uint32_t& c, uint32_t& j,
                                            uint32_t& nb) {
  c = b % chains;
  j = b / chains;
  nb = (grid - c + chains - 1) / chains;
}

// Keys the upfront kernel samples, and the count of one digit above which
// a byte is crowded.
constexpr int kCrowdSamples = 1024;
constexpr uint32_t kCrowdSample = kCrowdSamples / 16;

// Byte-0 histogram of one chain per workgroup plus OR/AND of the encoded
// keys; also clears the look-back descriptors (desc_words 32-bit words).
// Workgroups are split chain-major, so each flushes one 256-bin histogram.
// Workgroup 0 also reads 1024 keys evenly spaced over the input and sets bit
// b of red[3] when one digit of byte b holds more than kCrowdSample of them
// (a sixteenth): the pass before byte b then takes its uniform-digit launch
// (RANK 2), which counts byte b with one atomic per digit group, since one
// atomic per key would put most lanes of a wavefront on one LDS counter in
// turn.
template <int KT, int BLOCK, bool VEC>
__global__ __launch_bounds__(BLOCK) void upfront_kernel(const typename Codec<KT>::Bits* __restrict__ keys, uint32_t n,
                                                        uint32_t* __restrict__ hist, uint64_t* __restrict__ red,
                                                        uint32_t* __restrict__ desc, size_t desc_words, ChainMap cm,
                                                        int record_xcd, const uint32_t* __restrict__ gate) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int P = (int)sizeof(Bits);
  if (gated_off(gate)) return;
  __shared__ uint32_t h[kRadix];
  __shared__ Bits s_or[BLOCK / kWave], s_and[BLOCK / kWave];
  const int tid = threadIdx.x;
  for (int i = tid; i < kRadix; i += BLOCK) h[i] = 0;
  if (blockIdx.x == 0) {
    __shared__ uint32_t sh[P][kRadix];
    __shared__ uint32_t s_mask;
    for (int i = tid; i < P * kRadix; i += BLOCK) (&sh[0][0])[i] = 0;
    if (tid == 0) s_mask = 0;
    __syncthreads();
    for (int s = tid; s < kCrowdSamples; s += BLOCK) {
      const Bits e = Codec<KT>::enc(keys[((uint64_t)s * (n - 1)) / (kCrowdSamples - 1)]);
#pragma unroll
      for (int b = 0; b < P; ++b) atomicAdd(&sh[b][(uint32_t)(e >> (8 * b)) & 0xFFu], 1u);
    }
    __syncthreads();
    uint32_t m = 0;
    for (int i = tid; i < P * kRadix; i += BLOCK)
      if ((&sh[0][0])[i] > kCrowdSample) m |= 1u << (i / kRadix);
    if (m) atomicOr(&s_mask, m);
    __syncthreads();
    if (tid == 0 && s_mask) atomicOr(reinterpret_cast<unsigned long long*>(&red[3]), (unsigned long long)s_mask);
  }
  {
    const size_t gtid = (size_t)blockIdx.x * BLOCK + tid;
    const size_t gstride = (size_t)gridDim.x * BLOCK;
    uint4* d4 = reinterpret_cast<uint4*>(desc);
    const size_t n4 = desc_words / 4;
    for (size_t i = gtid; i < n4; i += gstride) d4[i] = make_uint4(0, 0, 0, 0);
    for (size_t i = n4 * 4 + gtid; i < desc_words; i += gstride) desc[i] = 0;
  }
  __syncthreads();
  uint32_t c, j, nb;
  block_chain(blockIdx.x, gridDim.x, cm.chains, c, j, nb);
  const size_t lo = (size_t)c * cm.chain_keys;
  const size_t hi = std::min<size_t>((size_t)n, lo + cm.chain_keys);
  Bits vor = 0, vand = ~Bits(0);
  // 8-byte keys count with the uniform-digit path: keys converted from
  // doubles (the generator's u64 keys) have their low 10-11 bits zero, so
  // almost every lane of every wavefront would take the same LDS counter in
  // turn (3.38 ms instead of 1.52 ms for the upfront read of 2^30 such keys);
  // for 4-byte keys the extra lane logic costs 5% and such keys are rare
  auto count = [&](Bits k) {
    const Bits e = Codec<KT>::enc(k);
    vor |= e;
    vand &= e;
    lds_count<sizeof(Bits) == 8>(h, (uint32_t)e & 0xFFu);
  };
  if constexpr (VEC) {
    // chain_keys is a multiple of the tile, so [lo, hi) starts 16-byte aligned
    constexpr int V = 16 / (int)sizeof(Bits);
    constexpr int U = 4;  // vectors in flight per thread
    using Vec = typename std::conditional<sizeof(Bits) == 4, uint4, ulonglong2>::type;
    const Vec* kv = reinterpret_cast<const Vec*>(keys + lo);
    const size_t nv = (hi - lo) / V;
    const size_t stride = (size_t)nb * BLOCK;
    size_t i = (size_t)j * BLOCK + tid;
    for (; i + (U - 1) * stride < nv; i += U * stride) {
      Vec v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = kv[i + u * stride];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if constexpr (sizeof(Bits) == 4) {
          count(v[u].x); count(v[u].y); count(v[u].z); count(v[u].w);
        } else {
          count(v[u].x); count(v[u].y);
        }
      }
    }
    for (; i < nv; i += stride) {
      const Vec v = kv[i];
      if constexpr (sizeof(Bits) == 4) {
        count(v.x); count(v.y); count(v.z); count(v.w);
      } else {
        count(v.x); count(v.y);
      }
    }
    for (size_t k = lo + nv * V + (size_t)j * BLOCK + tid; k < hi; k += stride) count(keys[k]);
  } else {
    for (size_t k = lo + (size_t)j * BLOCK + tid; k < hi; k += (size_t)nb * BLOCK) count(keys[k]);
  }
  vor = wave_or(vor);
  vand = wave_and(vand);
  if ((tid & (kWave - 1)) == 0) {
    s_or[tid / kWave] = vor;
    s_and[tid / kWave] = vand;
  }
  __syncthreads();
  if (tid == 0) {
    Bits a = 0, b = ~Bits(0);
    for (int i = 0; i < BLOCK / kWave; ++i) {
      a |= s_or[i];
      b &= s_and[i];
    }
    atomicOr(reinterpret_cast<unsigned long long*>(&red[0]), (unsigned long long)a);
    atomicAnd(reinterpret_cast<unsigned long long*>(&red[1]),
              (unsigned long long)b | (sizeof(Bits) == 4 ? 0xFFFFFFFF00000000ull : 0ull));
    if (record_xcd) atomicOr(reinterpret_cast<unsigned long long*>(&red[2]), 1ull << xcc_id());
  }
  for (int i = tid; i < kRadix; i += BLOCK) {
    const uint32_t v = h[i];
    if (v) atomicAdd(&hist[c * kRadix + i], v);
  }
}

// Clears `words` look-back descriptor words (between the rounds of a pass).
__global__ void clear_kernel(uint32_t* __restrict__ desc, size_t words, const uint32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  uint4* d4 = reinterpret_cast<uint4*>(desc);
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
  for (size_t i = g; i < words / 4; i += st) d4[i] = make_uint4(0, 0, 0, 0);
  for (size_t i = words / 4 * 4 + g; i < words; i += st) desc[i] = 0;
}

// Which bytes vary: bit b set iff byte b differs between keys.
__device__ __forceinline__ uint32_t varying_bytes(const uint64_t* red, int passes) {
  const uint64_t diff = red[0] ^ red[1];
  uint32_t m = 0;
  for (int b = 0; b < passes; ++b)
    if ((diff >> (8 * b)) & 0xFFu) m |= 1u << b;
  return m;
}

// Place of active pass k (of na) in the ping-pong sequence.
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

// Per-chain histogram of byte `b` of keys[0, n) (the input order), added
// into hist_b[chain][256]; workgroups split chain-major.
template <int KT, int BLOCK>
__device__ void histogram_only(const typename Codec<KT>::Bits* keys, uint32_t n, int b, uint32_t* hist_b,
                               const ChainMap& cm) {
  __shared__ uint32_t h[kRadix];
  for (int i = threadIdx.x; i < kRadix; i += BLOCK) h[i] = 0;
  __syncthreads();
  uint32_t c, j, nb;
  block_chain(blockIdx.x, gridDim.x, cm.chains, c, j, nb);
  const size_t lo = (size_t)c * cm.chain_keys;
  const size_t hi = std::min<size_t>((size_t)n, lo + cm.chain_keys);
  for (size_t i = lo + (size_t)j * BLOCK + threadIdx.x; i < hi; i += (size_t)nb * BLOCK)
    lds_count(h, digit_of<KT>(keys[i], 8 * b));
  __syncthreads();
  for (int i = threadIdx.x; i < kRadix; i += BLOCK)
    if (h[i]) atomicAdd(&hist_b[c * kRadix + i], h[i]);
}

template <int KT, typename V, int IPT>
__device__ __forceinline__ void load_tile(const typename Codec<KT>::Bits* __restrict__ src,
                                          const V* __restrict__ vsrc, uint32_t n, uint32_t wb, bool full,
                                          typename Codec<KT>::Bits (&k)[IPT],
                                          typename std::conditional<kHasVal<V>, V, uint32_t>::type (&v)[IPT]) {
  using Bits = typename Codec<KT>::Bits;
  if (full) {
#pragma unroll
    for (int i = 0; i < IPT; ++i) k[i] = src[wb + i * kWave];
    if constexpr (kHasVal<V>) {
#pragma unroll
      for (int i = 0; i < IPT; ++i) v[i] = vsrc[wb + i * kWave];
    }
  } else {
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const uint32_t idx = wb + i * kWave;
      k[i] = idx < n ? src[idx] : Bits(0);
      if constexpr (kHasVal<V>) v[i] = idx < n ? vsrc[idx] : V(0);
    }
  }
}

// Look-back window of one launch of a pass.  The descriptors of very large
// sorts would outgrow the scratch budget, so such a pass runs in rounds:
// round r covers tiles [first_t, first_t + tpr) of every chain, its
// descriptors occupy slots c * tpr + (tile index inside the round), and the
// chain's inclusive digit counts carry from one round to the next.  One round
// (first_t = 0, tpr = tiles per chain) is the plain pass.
struct PassRound {
  uint32_t first_t;          // r * tpr
  uint32_t tpr;              // tiles per chain and round
  int rounds;                // > 1: descriptors cleared before every launch
  const uint32_t* carry_in;  // [kChains][256] digit counts before the round (round > 0)
  uint32_t* carry_out;       // [kChains][256] digit counts after it (all rounds but the last)
};

// Tiles of chain c inside round rd.
__device__ __forceinline__ uint32_t round_tiles(const ChainMap& cm, const PassRound& rd, uint32_t c) {
  const uint32_t ct = min(cm.tpc, cm.tiles - c * cm.tpc);
  return ct > rd.first_t ? min(rd.tpr, ct - rd.first_t) : 0u;
}

// Next tile for a workgroup: tiles of its current chain in order, then of the
// chain `step` further on (modulo the chain count) once that one is used up.
// Thread-private state: the current chain and the number of chains of its
// group not yet found empty.
__device__ __forceinline__ uint32_t take_tile(uint32_t* __restrict__ cnt, const ChainMap& cm, const PassRound& rd,
                                              uint32_t& c, uint32_t& left, uint32_t step = 1) {
  while (left > 0) {
    const uint32_t t = atomicAdd(&cnt[c], 1u);
    if (t < round_tiles(cm, rd, c)) return c * cm.tpc + rd.first_t + t;
    c += step;
    if (c >= cm.chains) c -= cm.chains;
    --left;
  }
  return kNoTile;
}

// XCD-local look-back.  Each XCD has its own L2.  Measured on MI355X with a
// ping-pong between workgroups: a relaxed agent-scope load sees a
// workgroup-scope store of another workgroup on the same XCD (the store stays
// in that L2, the load is served from it: 0.56 us round trip, against
// 0.99 us when the store is agent-scope as well), but a workgroup on another
// XCD does not see it before the kernel ends.  So when the upfront kernel has
// run on all eight XCDs (ids 0..7) and the chain count is a multiple of
// eight, chain c is served only by workgroups running on XCD c mod 8, and
// its descriptors are stored with workgroup scope; every tile of a chain is
// then written and read through one L2.  Otherwise descriptors are stored
// with agent scope and any workgroup may serve any chain.
template <typename D>
__device__ __forceinline__ void publish(D* p, D v, bool xcd_local) {
  if (xcd_local)
    st_wg(p, v);
  else
    st_agent(p, v);
}

// ---------------------------------------------------------------------------
// One onesweep pass over byte `pass`, persistent workgroups, chained tiles.
// hist is [P][kChains][256]; counters is [P][kChains].
// ---------------------------------------------------------------------------
// Digit base of a chain: `dbase` (keys with a smaller digit) plus the keys of
// earlier chains with the same digit.  The row offset is made opaque so the
// compiler does not hold the per-chain addresses in registers across the
// tile loop (this runs only when a workgroup changes chain); the loads are
// unconditional (every chain row exists) so that all are in flight at once.
__device__ __forceinline__ uint32_t chain_digit_base(const uint32_t* __restrict__ hp, uint32_t dbase, uint32_t chain,
                                                     int tid) {
  uint32_t o = (uint32_t)tid;
  asm volatile("" : "+v"(o));
  uint32_t h[kChains];
#pragma unroll
  for (uint32_t c = 0; c < (uint32_t)kChains; ++c) h[c] = hp[c * kRadix + o];
  uint32_t b = dbase;
#pragma unroll
  for (uint32_t c = 0; c < (uint32_t)kChains; ++c) b += c < chain ? h[c] : 0u;
  return b;
}

// RANK selects how each wavefront ranks its keys' digits in LDS:
//   0  one atomic per key (ds_add_rtn_u32 on the wavefront's counter of the
//      digit; the only variant launched for smaller sorts);
//   1  the same, first of a pair of launches for one pass: exits at once
//      when the pass is skewed;
//   2  uniform-digit ranking (the lanes of a wavefront that share the first
//      lane's digit take one LDS atomic instead of one each), second of the
//      pair: exits at once unless the pass is skewed;
//   3  peer-mask ranking (rank_wave), chosen at run time with
//      MGS_RANK_PEERS=1 for 4-byte keys (single-round passes).
// Variants 0-2 are stable because the LDS serves the conflicting lanes of one
// ds_add_rtn_u32 in ascending lane order: measured on gfx950 (and re-checked
// by every multi-pass parity test with payloads), not a documented guarantee.
// Variant 3 relies only on the in-order execution of one wavefront's DS
// instructions, at the cost of an OR, a read and a clear per key.
// A pass is skewed when one digit of its byte holds more than
// n >> MGS_SKEW_SHIFT of the keys (heavy duplicates).  The choice is made by
// the pair of launches, not by a branch inside one kernel, because a kernel
// holding both ranking paths needs more registers than one workgroup of 1024
// threads per CU can give and spills.
//
// Per tile, waves 4..7 (threads 256..511, one per digit) own the look-back:
// they publish the tile's digit counts, take the next tile and issue the
// first look-back window while waves 0..3 scan the counts into tile-local
// offsets; they take one more look-back round before the scatter and finish
// after it, so the look-back latency overlaps the scan and the scatter.
// Every wavefront loads its part of the next tile as soon as its keys of the
// current tile are in LDS.
// MULTI: the pass runs in rounds (descriptor budget exceeded): tile numbers
// inside a round and the digit counts carried between rounds.  Single-round
// passes compile that bookkeeping out (measured 5% per pass on 2^29 u64 keys).
template <int KT, typename V, int BLOCK, int IPT, typename D, int WAVES, int RANK = 0, bool MULTI = false>
__global__ __launch_bounds__(BLOCK, WAVES) void pass_kernel(
    const typename Codec<KT>::Bits* k_in, typename Codec<KT>::Bits* k_alt, typename Codec<KT>::Bits* k_out,
    const V* v_in, V* v_alt, V* v_out, uint32_t n, int pass, int aliased, ChainMap cm,
    const uint64_t* __restrict__ red, uint32_t* __restrict__ hist, uint32_t* __restrict__ counters,
    D* __restrict__ desc, uint32_t* __restrict__ status, uint32_t spin_limit, const uint32_t* __restrict__ gate,
    PassRound rd_arg) {
  PassRound rd = rd_arg;
  if constexpr (!MULTI) {
    rd.first_t = 0;
    rd.rounds = 1;
    rd.carry_in = nullptr;
    rd.carry_out = nullptr;
  }
  using Bits = typename Codec<KT>::Bits;
  constexpr int P = (int)sizeof(Bits);
  constexpr int W = BLOCK / kWave;
  constexpr int TILE = BLOCK * IPT;
  constexpr bool HV = kHasVal<V>;
  using VS = typename std::conditional<HV, V, uint32_t>::type;
  static_assert(W >= 8, "the look-back wavefronts 4..7 need >= 512 threads");

  constexpr bool kUni = RANK == 2;
  constexpr bool kPeers = RANK == 3;
  if (gated_off(gate)) return;
  const uint32_t vary = varying_bytes(red, P);
  if (!(vary & (1u << pass))) {
    if constexpr (kUni) return;  // the RANK 1 launch of the pair does this
    // Trivial byte.  If it is the last trivial byte below the first varying
    // one, nobody has counted that byte yet: do it here from the input.
    const uint32_t below = vary & ((2u << pass) - 1u);
    if (below == 0 && pass + 1 < P && (vary & (1u << (pass + 1))))
      histogram_only<KT, BLOCK>(k_in, n, pass + 1, hist + (size_t)(pass + 1) * kChains * kRadix, cm);
    return;
  }
  const int k = __builtin_popcount(vary & ((1u << pass) - 1u));
  const int na = __builtin_popcount(vary);
  const uint32_t above = vary & ~((2u << pass) - 1u);
  const int next = above ? __builtin_ctz(above) : -1;
  const bool next_crowded = next >= 0 && ((red[3] >> next) & 1u) != 0;
  const int sb = src_of(k, na, aliased), db = dst_of(k, na, aliased);
  const Bits* __restrict__ src = pick_src(sb, k_in, (const Bits*)k_alt, (const Bits*)k_out);
  Bits* __restrict__ dst = pick_dst(db, k_alt, k_out);
  const V* __restrict__ vsrc = nullptr;
  V* __restrict__ vdst = nullptr;
  if constexpr (HV) {
    vsrc = pick_src(sb, v_in, (const V*)v_alt, (const V*)v_out);
    vdst = pick_dst(db, v_alt, v_out);
  }
  // status rotation by pass (descriptors cleared once per sort); with rounds
  // the descriptors are cleared before every launch instead
  const int shift = 8 * pass, parity = rd.rounds > 1 ? 0 : (k & 1);
  uint32_t* __restrict__ cnt = counters + ((size_t)pass * rd.rounds + rd.first_t / rd.tpr) * kChains;
  const uint32_t* __restrict__ hp = hist + (size_t)pass * kChains * kRadix;

  __shared__ Bits keys[TILE];
  __shared__ VS vstage[HV ? TILE : 1];
  __shared__ uint64_t peers[kPeers ? W : 1][kRadix];
  __shared__ uint32_t off[W][kRadix];
  __shared__ uint32_t nh[kChains][kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint32_t s_prefix[kRadix];
  __shared__ uint32_t sums[4];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_skew;

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;
  const bool lb_wave = tid >= kRadix && tid < 2 * kRadix;
  // Tiles are taken by the first look-back thread (its wavefront issues no
  // descriptor stores before the atomic); its acquisition state:
  const int taker = kRadix;
  const bool xcd_local = MGS_XCD_LOCAL && cm.chains % kXcds == 0 && ld_agent(&red[2]) == (1ull << kXcds) - 1;
  uint32_t my_chain = blockIdx.x % cm.chains, chains_left = cm.chains, chain_step = 1;
  if (xcd_local) {
    const uint32_t g = cm.chains / kXcds;
    my_chain = xcc_id() + kXcds * ((blockIdx.x / kXcds) % g);
    chains_left = g;
    chain_step = kXcds;
  }

  // Digit base of this byte (keys with a smaller digit), held by the thread
  // of each digit; the base of a chain adds the keys of earlier chains with
  // the same digit, recomputed when the workgroup moves to another chain.
  uint32_t dbase = 0, cur_chain = 0xFFFFFFFFu, cur_base = 0;
  {
    uint32_t tot = 0;
    if (tid < kRadix) {
#pragma unroll
      for (uint32_t c = 0; c < (uint32_t)kChains; ++c)
        if (c < cm.chains) tot += hp[c * kRadix + tid];  // independent loads, all in flight
    }
    if constexpr (RANK == 1 || RANK == 2) {
      if (tid == 0) s_skew = 0;
    }
    dbase = block_excl_scan_256(tot, sums, tid);  // contains a barrier
    if constexpr (RANK == 1 || RANK == 2) {
      // a crowded next byte also takes the uniform-digit launch, whose
      // write-out counts that byte with one atomic per digit group
      if ((tid < kRadix && tot > (n >> MGS_SKEW_SHIFT)) || (tid == 0 && next_crowded)) s_skew = 1;
    }
    if (tid < kRadix)
      for (uint32_t c = 0; c < (uint32_t)kChains; ++c) nh[c][tid] = 0;
    for (int i = tid; i < W * kRadix; i += BLOCK) (&off[0][0])[i] = 0;
    if constexpr (kPeers) {
      for (int i = tid; i < W * kRadix; i += BLOCK) (&peers[0][0])[i] = 0;
    }
    if constexpr (RANK == 1 || RANK == 2) {
      // Not this launch's pass: leave before taking a tile (the whole
      // workgroup reads the same flag, so it leaves together).
      __syncthreads();
      if ((s_skew != 0) != kUni) return;
    }
    if (tid == taker) s_tile = take_tile(cnt, cm, rd, my_chain, chains_left, chain_step);
    __syncthreads();
  }
  uint32_t tile = s_tile;
  Bits kc[IPT];
  VS vc[IPT];
  if (tile != kNoTile) {
    const uint32_t tb = tile * (uint32_t)TILE;
    load_tile<KT, V, IPT>(src, vsrc, n, tb + (uint32_t)(w * IPT * kWave + lane), n - tb >= (uint32_t)TILE, kc, vc);
  }

  while (tile != kNoTile) {
    const uint32_t chain = tile / cm.tpc;
    const uint32_t lt = tile - chain * cm.tpc - rd.first_t;  // tile index inside this round of its chain
    const uint32_t tile_base = tile * (uint32_t)TILE;
    const uint32_t tile_cnt = min((uint32_t)TILE, n - tile_base);
    const bool full = tile_cnt == (uint32_t)TILE;
    const uint32_t wb = tile_base + (uint32_t)(w * IPT * kWave + lane);
    // per item: digit << 16 | rank among the wavefront's earlier items with
    // that digit (< IPT * 64)
    uint32_t pk[IPT];
    uint32_t* ct = off[w];
    if (full) {
      if constexpr (kPeers) {
#pragma unroll
        for (int i = 0; i < IPT; ++i) {
          const uint32_t d = digit_of<KT>(kc[i], shift);
          pk[i] = (d << 16) | rank_wave(d, peers[w], ct);
        }
      } else {
        uint32_t dg[IPT];
#pragma unroll
        for (int i = 0; i < IPT; ++i) dg[i] = digit_of<KT>(kc[i], shift);
        rank_items_full<IPT, kUni>(dg, ct, pk);
      }
    } else {
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint32_t d = digit_of<KT>(kc[i], shift);
        if (wb + i * kWave < n) {
          if constexpr (kPeers)
            pk[i] = (d << 16) | rank_wave(d, peers[w], ct);
          else
            pk[i] = (d << 16) | rank_wave_atomic<false>(d, ct);
        } else {
          pk[i] = 0u;
        }
      }
    }
    // The keys are in registers now; tell the compiler so, so that the
    // scatter below does not wait on memory operations issued meanwhile.
#pragma unroll
    for (int i = 0; i < IPT; ++i) asm volatile("" : "+v"(kc[i]));
    __syncthreads();

    uint32_t total = 0;
    D* const chain_desc = desc + (size_t)chain * rd.tpr * kRadix;
    D win[kPassWindow];
    uint32_t taken = kNoTile;
    if (lb_wave) {
      const uint32_t d = (uint32_t)tid - kRadix;
      if (tid == taker) taken = take_tile(cnt, cm, rd, my_chain, chains_left, chain_step);
      // the round's first tile of a chain starts from the counts the
      // previous round carried over
      const uint32_t carry = (lt == 0 && rd.carry_in != nullptr) ? rd.carry_in[chain * kRadix + d] : 0u;
#pragma unroll
      for (int i = 0; i < W; ++i) total += off[i][d];
      publish(chain_desc + (size_t)lt * kRadix + d,
              Desc<D>::make(lt == 0 ? kStInclusive : kStAggregate, (D)(carry + total), parity), xcd_local);
      if (lt > 0) lookback_load(chain_desc, (int64_t)lt - 1, d, win);
    } else if (tid < kRadix) {
#pragma unroll
      for (int i = 0; i < W; ++i) total += off[i][tid];
    }
    const uint32_t excl = block_excl_scan_256(tid < kRadix ? total : 0u, sums, tid);
    if (tid < kRadix) {
      uint32_t run = excl;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const uint32_t c = off[i][tid];
        off[i][tid] = run;
        run += c;
      }
      if (chain != cur_chain) {
        cur_chain = chain;
        cur_base = chain_digit_base(hp, dbase, chain, tid);
      }
      s_prefix[tid] = cur_base - excl;  // the digit base without the look-back prefix
    }
    if (tid == taker) s_tile = taken;
    __syncthreads();
    int64_t lb_t = (int64_t)lt - 1;
    uint32_t lb_prefix = 0;
    bool lb_done = false;
    if (lb_wave && lt > 0) lb_done = lookback_step(chain_desc, lb_t, (uint32_t)tid - kRadix, parity, win, lb_prefix);

    if (full) {
      // all offset reads first, then all writes (branch-free; the clamp only
      // keeps a corrupted offset inside the staging buffer)
      uint32_t pos[IPT];
#pragma unroll
      for (int i = 0; i < IPT; ++i) pos[i] = min(off[w][pk[i] >> 16] + (pk[i] & 0xFFFFu), (uint32_t)TILE - 1);
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        keys[pos[i]] = kc[i];
        if constexpr (HV) vstage[pos[i]] = vc[i];
      }
    } else {
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint32_t pos = off[w][pk[i] >> 16] + (pk[i] & 0xFFFFu);
        if (wb + i * kWave < n && pos < (uint32_t)TILE) {
          keys[pos] = kc[i];
          if constexpr (HV) vstage[pos] = vc[i];
        }
      }
    }
    // This wavefront's rank counters and keys are free: clear the counters
    // and load its part of the next tile.
    Bits kn[IPT];
    VS vn[IPT];
    const uint32_t nxt = s_tile;
    if (nxt != kNoTile) {
      const uint32_t tb = nxt * (uint32_t)TILE;
      load_tile<KT, V, IPT>(src, vsrc, n, tb + (uint32_t)(w * IPT * kWave + lane), n - tb >= (uint32_t)TILE, kn, vn);
    }
    for (int i = lane; i < kRadix; i += kWave) off[w][i] = 0;
    if (lb_wave) {
      // this wavefront's part of the tile is in LDS: finish the look-back
      const uint32_t d = (uint32_t)tid - kRadix;
      uint32_t prefix = 0;
      if (lt == 0 && rd.carry_in != nullptr) prefix = rd.carry_in[chain * kRadix + d];
      if (lt > 0) {
        bool gave_up = false;
        prefix = lb_done ? lb_prefix : lookback_from(chain_desc, lb_t, d, parity, win, lb_prefix, spin_limit, gave_up);
        if (gave_up) atomicOr(status, kStatusLookback);
        publish(chain_desc + (size_t)lt * kRadix + d, Desc<D>::make(kStInclusive, (D)(prefix + total), parity),
                xcd_local);
      }
      gbase[d] = s_prefix[d] + prefix;
      if (rd.carry_out != nullptr && lt + 1 == round_tiles(cm, rd, chain)) rd.carry_out[chain * kRadix + d] = prefix + total;
    }
    __syncthreads();

    // Write-out: consecutive threads store consecutive positions of each
    // digit run.  o < n for every key of a well-formed pass; the clamp only
    // keeps a corrupted prefix in bounds and leaves the stores unconditional.
    // On the way out each key is counted in the next varying byte's
    // histogram of the chain its output position falls in.
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = (uint32_t)(j * BLOCK + tid);
      if (full || i < tile_cnt) {
        const Bits kk = keys[i];
        const uint32_t o = min(gbase[digit_of<KT>(kk, shift)] + i, n - 1);
        dst[o] = kk;
        if constexpr (HV) vdst[o] = vstage[i];
        if (next >= 0) lds_count<kUni>(&nh[0][0], chain_of(o, cm) * kRadix + digit_of<KT>(kk, 8 * next));
      }
    }
    __syncthreads();
    tile = nxt;
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      kc[i] = kn[i];
      if constexpr (HV) vc[i] = vn[i];
    }
  }
  if (next >= 0 && tid < kRadix) {
    uint32_t* hn = hist + (size_t)next * kChains * kRadix;
    for (uint32_t c = 0; c < cm.chains; ++c)
      if (nh[c][tid]) atomicAdd(&hn[c * kRadix + tid], nh[c][tid]);
  }
}

// Copies the result into the output buffers when the passes left it in the
// scratch buffer (or did not run), and checks that every varying byte's pass
// took every tile of every chain (a chain no workgroup served -- say, an XCD
// without resident workgroups in XCD-local mode -- would leave its keys
// unwritten): take_tile() counts every attempt, so a chain's counter ends at
// or above its tile count iff all of its tiles were taken.
template <typename Bits, typename V>
__global__ void fixup_kernel(const Bits* k_in, const Bits* k_alt, Bits* k_out, const V* v_in, const V* v_alt, V* v_out,
                             uint32_t n, const uint64_t* __restrict__ red, int aliased,
                             const uint32_t* __restrict__ counters, ChainMap cm, uint32_t tpr, int rounds,
                             uint32_t* __restrict__ status, const uint32_t* __restrict__ gate) {
  constexpr int P = (int)sizeof(Bits);
  if (gated_off(gate)) return;
  const uint32_t vary = varying_bytes(red, P);
  if (blockIdx.x == 0) {
    for (uint32_t i = threadIdx.x; i < (uint32_t)(P * rounds * kChains); i += blockDim.x) {
      const uint32_t c = i % kChains, r = (i / kChains) % rounds, p = i / (kChains * rounds);
      if ((vary >> p) & 1u && c < cm.chains) {
        const PassRound rd{r * tpr, tpr, rounds, nullptr, nullptr};
        if (counters[i] < round_tiles(cm, rd, c)) atomicOr(status, kStatusChainUnserved);
      }
    }
  }
  const int na = __builtin_popcount(vary);
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

// XCD-local chains can be switched off at run time with MGS_XCD_LOCAL=0 in
// the environment (read on every call): the upfront kernel then records no
// XCDs and every pass uses agent-scope descriptors.
int xcd_local_allowed() {
  const char* e = std::getenv("MGS_XCD_LOCAL");
  return (e != nullptr && e[0] == '0' && e[1] == 0) ? 0 : 1;
}

// Smallest sort that launches each pass as a pair of ranking variants
// (MGS_SKEW_PAIR_MIN keys; the environment variable of the same name, a
// decimal key count, overrides it).
size_t skew_pair_min() {
  const char* e = std::getenv("MGS_SKEW_PAIR_MIN");
  if (e != nullptr && e[0] != 0) {
    char* end = nullptr;
    const unsigned long long v = std::strtoull(e, &end, 10);
    if (end != nullptr && *end == 0) return (size_t)v;
  }
  return (size_t)MGS_SKEW_PAIR_MIN;
}

// Look-back wait bound of the pass kernels (MGS_LB_SPIN_LIMIT sleeps; the
// environment variable of that name, a decimal count, overrides it, so a
// test can make every wait give up and check that the status word says so).
uint32_t lookback_spin_limit() {
  const char* e = std::getenv("MGS_LB_SPIN_LIMIT");
  if (e != nullptr && e[0] != 0) {
    char* end = nullptr;
    const unsigned long long v = std::strtoull(e, &end, 10);
    if (end != nullptr && *end == 0) return (uint32_t)std::min<unsigned long long>(v, 0xFFFFFFFFull);
  }
  return kLookbackSpinLimit;
}

// Peer-mask ranking (pass_kernel RANK 3) for 4-byte keys when the
// environment holds MGS_RANK_PEERS=1.
bool rank_peers_requested() {
  const char* e = std::getenv("MGS_RANK_PEERS");
  return e != nullptr && e[0] == '1' && e[1] == 0;
}

// Look-back descriptor memory of one sort: at most this many bytes
// (MGS_DESC_BUDGET, a decimal byte count in the environment, overrides it);
// a sort whose descriptors would need more runs each pass in rounds.
size_t desc_budget() {
  const char* e = std::getenv("MGS_DESC_BUDGET");
  if (e != nullptr && e[0] != 0) {
    char* end = nullptr;
    const unsigned long long v = std::strtoull(e, &end, 10);
    if (end != nullptr && *end == 0 && v >= 2 * kChains * kRadix * 4) return (size_t)v;
  }
  return size_t(96) << 20;
}

struct SortLayout {
  size_t status, alt_k, alt_v, hist, red, counters, carry, msd_ctl, desc, total;
  uint32_t num_tiles;
  uint32_t rounds, tpr;  // rounds per pass and tiles per chain and round
  bool wide_desc;
};

template <int KB, int VB, bool BIG>
SortLayout sort_layout_cfg(size_t n) {
  constexpr size_t tile = (size_t)PassCfg<KB, VB, BIG>::kTile;
  SortLayout L{};
  L.num_tiles = (uint32_t)((n + tile - 1) / tile);
  const ChainMap cm = make_chain_map(n, (uint32_t)tile, kChains);
  // descriptor values are prefixes inside one chain: 30 bits hold them
  // unless a chain reaches 2^30 keys
  L.wide_desc = cm.chain_keys >= (size_t(1) << 30);
  const size_t dbytes = L.wide_desc ? 8 : 4;
  const size_t full = (size_t)cm.chains * cm.tpc * kRadix * dbytes;
  const size_t budget = desc_budget();
  L.rounds = 1;
  L.tpr = cm.tpc;
  size_t desc_bytes = full;
  if (full > budget) {
    // the fewest rounds whose window fits; the full budget is reserved, so
    // that the scratch size never decreases with n
    uint32_t r = (uint32_t)((full + budget - 1) / budget);
    while ((size_t)cm.chains * ((cm.tpc + r - 1) / r) * kRadix * dbytes > budget) ++r;
    L.rounds = r;
    L.tpr = (cm.tpc + r - 1) / r;
    desc_bytes = budget;
  }
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align_up(o + bytes);
    return at;
  };
  L.status = take(sizeof(uint32_t));  // first word of temp: the call's status word
  L.alt_k = take(n * KB);
  L.alt_v = take(n * VB);
  L.hist = take((size_t)KB * kChains * kRadix * 4);
  L.red = take(4 * sizeof(uint64_t));
  L.counters = take((size_t)KB * L.rounds * kChains * 4);
  L.carry = take(2 * (size_t)kChains * kRadix * 4);
  L.msd_ctl = take(sizeof(MsdCtl));  // the three-pass path's decision word: read by the LSD kernels
  L.desc = take(desc_bytes);
  L.total = o;
  return L;
}

template <int KB, int VB>
SortLayout sort_layout(size_t n) {
  return use_big_tiles(KB, VB, n) ? sort_layout_cfg<KB, VB, true>(n) : sort_layout_cfg<KB, VB, false>(n);
}

// Scratch bytes for n keys: the layout's size, made non-decreasing in n.
// Wide records switch to larger tiles (fewer descriptors) at kBigTileMin
// keys, so just above it the small-tile layout of kBigTileMin - 1 keys can
// be the larger one; a buffer sized for n must serve every smaller sort.
template <int KB, int VB>
size_t sort_bytes(size_t n) {
  const SortLayout L = sort_layout<KB, VB>(n);
  size_t b = L.total;
  // the three-pass path's region wherever it may run, and from its lowest
  // default size on whatever the size ranges, so that the answer does not
  // decrease with n
  if (KB == 4 && VB == 0 && (msd_eligible(n) || n >= MsdCfg::kMinKeys))
    b = std::max(b, msd_layout(std::min(n, MsdCfg::kMaxKeys), L.desc, L.msd_ctl).end);
  if (KB + VB > 4 && n >= kBigTileMin) b = std::max(b, sort_layout_cfg<KB, VB, false>(kBigTileMin - 1).total);
  return b;
}

// Workgroups of `kernel` that fit on the current device at once (persistent
// grid).  Cached per (kernel, device), so that calls made while a stream is
// being captured into a graph issue no device queries.
template <typename K>
uint32_t resident_blocks(K kernel, int block, size_t dyn_lds) {
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
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, dyn_lds) != hipSuccess) return 0;
  const uint32_t r = (uint32_t)std::max(1, per_cu) * (uint32_t)std::max(1, cus);
  std::lock_guard<std::mutex> lock(mu);
  cache[key] = r;
  return r;
}

template <int KT, int VT, bool BIG>
int sort_typed_cfg(const void* kin, void* kout, const void* vin, void* vout, size_t n, void* temp, size_t temp_bytes,
                   hipStream_t s, Marks& mark) {
  using Bits = typename Codec<KT>::Bits;
  using V = typename ValOf<VT>::T;
  constexpr int KB = sizeof(Bits);
  constexpr int VB = kValBytes<V>;
  constexpr int P = KB;
  using Cfg = PassCfg<KB, VB, BIG>;
  using SCfg = SmallCfg<KB, VB>;
  using VP = typename std::conditional<kHasVal<V>, V, uint32_t>::type;

  const Bits* ki = static_cast<const Bits*>(kin);
  Bits* ko = static_cast<Bits*>(kout);
  const V* vi = static_cast<const V*>(vin);
  V* vo = static_cast<V*>(vout);

  if (n <= (size_t)SCfg::kTile) {
    mark(s, "small_sort");
    small_sort_kernel<KT, V, SCfg::kBlock, SCfg::kIpt><<<1, SCfg::kBlock, 0, s>>>(ki, ko, vi, vo, (uint32_t)n);
    mark(s);
    return check_launch();
  }

  const SortLayout L = sort_layout_cfg<KB, VB, BIG>(n);
  // 4-byte keys alone: the three-pass path first (its scratch overlaps the
  // LSD descriptors, which only a sort it turned itself off for uses)
  const bool msd = KB == 4 && VB == 0 && msd_eligible(n);
  const MsdLayout ML = msd_layout(n, L.desc, L.msd_ctl);
  const size_t need = msd ? std::max(L.total, ML.end) : L.total;
  if (temp == nullptr || temp_bytes < need) return MGS_ERR_TEMP_TOO_SMALL;
  char* t = static_cast<char*>(temp);
  uint32_t* status = reinterpret_cast<uint32_t*>(t + L.status);
  Bits* ka = reinterpret_cast<Bits*>(t + L.alt_k);
  V* va = reinterpret_cast<V*>(reinterpret_cast<VP*>(t + L.alt_v));
  uint32_t* hist = reinterpret_cast<uint32_t*>(t + L.hist);
  uint64_t* red = reinterpret_cast<uint64_t*>(t + L.red);
  uint32_t* counters = reinterpret_cast<uint32_t*>(t + L.counters);
  void* desc = t + L.desc;
  const int aliased = (kin == kout) || (kHasVal<V> && vin == vout);
  const ChainMap cm = make_chain_map(n, (uint32_t)Cfg::kTile, kChains);
  // descriptor slots in use: one per tile, or one per tile of a round
  const size_t desc_slots = L.rounds > 1 ? (size_t)cm.chains * L.tpr : (size_t)L.num_tiles;
  const size_t desc_words = desc_slots * kRadix * (L.wide_desc ? 2 : 1);
  const uint32_t spin_limit = lookback_spin_limit();
  const uint32_t* gate = nullptr;
  if constexpr (KB == 4 && VB == 0) {
    if (msd) {
      const int rc = msd_launch<KT>(kin, kout, ka, n, t, ML, status, spin_limit, s, mark);
      if (rc != MGS_OK) return rc;
      gate = reinterpret_cast<const uint32_t*>(t + ML.ctl);  // MsdCtl::off
    }
  }

  mark(s, "init");
  init_kernel<<<1, 1024, 0, s>>>(hist, (uint32_t)(P * kChains * kRadix), counters, (uint32_t)(P * L.rounds * kChains),
                                 red, status, gate);
  mark(s, "upfront");
  {
    constexpr int HB = 256;
    // chain-major split: the same number of workgroups per chain
    const uint32_t per_chain = (uint32_t)std::min<size_t>(
        std::max<size_t>((cm.chain_keys + 16 * HB - 1) / (16 * HB), 1), std::max<uint32_t>(2048 / cm.chains, 1));
    const unsigned grid = per_chain * cm.chains;
    if ((reinterpret_cast<uintptr_t>(kin) % 16) == 0)
      upfront_kernel<KT, HB, true><<<grid, HB, 0, s>>>(ki, (uint32_t)n, hist, red, (uint32_t*)desc, desc_words, cm,
                                                       xcd_local_allowed(), gate);
    else
      upfront_kernel<KT, HB, false><<<grid, HB, 0, s>>>(ki, (uint32_t)n, hist, red, (uint32_t*)desc, desc_words, cm,
                                                        xcd_local_allowed(), gate);
  }
  uint32_t* carry = reinterpret_cast<uint32_t*>(t + L.carry);
  PassRound rd{0, L.tpr, (int)L.rounds, nullptr, nullptr};
  auto launch_pass = [&](auto kernel, auto* dsc, int p) {
    const uint32_t res = resident_blocks(kernel, Cfg::kBlock, 0);
    uint32_t grid = std::min<uint32_t>(L.num_tiles, res ? res : L.num_tiles);
    grid = std::max<uint32_t>(grid, cm.chains);  // every chain has a home workgroup
    kernel<<<grid, Cfg::kBlock, 0, s>>>(ki, ka, ko, vi, va, vo, (uint32_t)n, p, aliased, cm, red, hist, counters,
                                        dsc, status, spin_limit, gate, rd);
  };
  // Large sorts launch each pass as a pair of pass_kernel variants, one-
  // atomic-per-key ranking and uniform-digit ranking; the variant that does
  // not suit the pass's digit counts exits at once (a few microseconds).
  const bool rank_pair = BIG && n >= skew_pair_min();
  const bool peers = KB + VB == 4 && rank_peers_requested();
  auto launch_ranked = [&](auto* dsc, int p) {
    using DT = std::remove_pointer_t<decltype(dsc)>;
    if (L.rounds > 1) {
      // passes in rounds (beyond the descriptor budget: about 1.4e9 keys and
      // more): one atomic per key, whatever the byte's skew
      launch_pass(pass_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, DT, Cfg::kWavesPerEu, 0, true>, dsc, p);
      return;
    }
    if constexpr (KB + VB == 4) {
      if (peers) {
        launch_pass(pass_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, DT, Cfg::kWavesPerEu, 3>, dsc, p);
        return;
      }
    }
    if constexpr (BIG) {
      if (rank_pair) {
        launch_pass(pass_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, DT, Cfg::kWavesPerEu, 1>, dsc, p);
        launch_pass(pass_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, DT, Cfg::kWavesPerEu, 2>, dsc, p);
        return;
      }
    }
    launch_pass(pass_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, DT, Cfg::kWavesPerEu, 0>, dsc, p);
  };
  static const char* const kPassName[kMaxPasses] = {"pass0", "pass1", "pass2", "pass3",
                                                    "pass4", "pass5", "pass6", "pass7"};
  for (int p = 0; p < P; ++p) {
    mark(s, kPassName[p]);
    for (uint32_t r = 0; r < L.rounds; ++r) {
      if (L.rounds > 1) {
        // every round starts from cleared descriptors (the first round of
        // the first pass finds them cleared by the upfront kernel)
        if (p > 0 || r > 0) clear_kernel<<<1024, 256, 0, s>>>(static_cast<uint32_t*>(desc), desc_words, gate);
        rd.first_t = r * L.tpr;
        rd.carry_in = r > 0 ? carry + ((r - 1) & 1) * kChains * kRadix : nullptr;
        rd.carry_out = r + 1 < L.rounds ? carry + (r & 1) * kChains * kRadix : nullptr;
      }
      if (L.wide_desc)
        launch_ranked(static_cast<uint64_t*>(desc), p);
      else
        launch_ranked(static_cast<uint32_t*>(desc), p);
    }
  }
  mark(s, "fixup");
  {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    fixup_kernel<Bits, V><<<grid, 256, 0, s>>>(ki, ka, ko, vi, va, vo, (uint32_t)n, red, aliased, counters, cm, L.tpr,
                                               (int)L.rounds, status, gate);
  }
  mark(s);
  return check_launch();
}

template <int KT, int VT>
int sort_typed(const void* kin, void* kout, const void* vin, void* vout, size_t n, void* temp, size_t temp_bytes,
               hipStream_t s, Marks& mark) {
  constexpr int KB = (int)sizeof(typename Codec<KT>::Bits);
  constexpr int VB = kValBytes<typename ValOf<VT>::T>;
  if constexpr (KB + VB <= 4)
    return sort_typed_cfg<KT, VT, true>(kin, kout, vin, vout, n, temp, temp_bytes, s, mark);
  else if (use_big_tiles(KB, VB, n))
    return sort_typed_cfg<KT, VT, true>(kin, kout, vin, vout, n, temp, temp_bytes, s, mark);
  else
    return sort_typed_cfg<KT, VT, false>(kin, kout, vin, vout, n, temp, temp_bytes, s, mark);
}

template <int KT>
int sort_by_val(mgs_val vt, const void* kin, void* kout, const void* vin, void* vout, size_t n, void* temp,
                size_t temp_bytes, hipStream_t s, Marks& m) {
  switch (vt) {
    case MGS_VAL_NONE: return sort_typed<KT, MGS_VAL_NONE>(kin, kout, vin, vout, n, temp, temp_bytes, s, m);
    case MGS_VAL_U32: return sort_typed<KT, MGS_VAL_U32>(kin, kout, vin, vout, n, temp, temp_bytes, s, m);
    case MGS_VAL_U64: return sort_typed<KT, MGS_VAL_U64>(kin, kout, vin, vout, n, temp, temp_bytes, s, m);
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
    if (vb == 0) return sort_bytes<4, 0>(n);
    if (vb == 4) return sort_bytes<4, 4>(n);
    return sort_bytes<4, 8>(n);
  }
  if (vb == 0) return sort_bytes<8, 0>(n);
  if (vb == 4) return sort_bytes<8, 4>(n);
  return sort_bytes<8, 8>(n);
}

namespace {
thread_local std::string g_stage_names;

template <int KB, int VB>
int path_of(size_t n, const void* temp, size_t temp_bytes, hipStream_t s) {
  using SCfg = SmallCfg<KB, VB>;
  if (n <= (size_t)SCfg::kTile) return MGS_PATH_SMALL;
  if (!(KB == 4 && VB == 0 && msd_eligible(n))) return MGS_PATH_LSD;
  const SortLayout L = sort_layout<KB, VB>(n);
  const MsdLayout ML = msd_layout(n, L.desc, L.msd_ctl);
  if (temp == nullptr || temp_bytes < ML.end) return MGS_ERR_TEMP_TOO_SMALL;
  uint32_t off = 0;
  hipError_t e = hipMemcpyAsync(&off, static_cast<const char*>(temp) + ML.ctl, sizeof(off), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return record_hip(e);
  return off == 0 ? MGS_PATH_THREE_PASS : MGS_PATH_LSD;
}
}  // namespace

int sort_path(mgs_key kt, mgs_val vt, size_t n, const void* temp, size_t temp_bytes, hipStream_t s) {
  const int kb = key_bytes(kt), vb = val_bytes(vt);
  if (kb <= 0 || vb < 0) return MGS_ERR_BAD_TYPE;
  if (kb == 4) {
    if (vb == 0) return path_of<4, 0>(n, temp, temp_bytes, s);
    if (vb == 4) return path_of<4, 4>(n, temp, temp_bytes, s);
    return path_of<4, 8>(n, temp, temp_bytes, s);
  }
  if (vb == 0) return path_of<8, 0>(n, temp, temp_bytes, s);
  if (vb == 4) return path_of<8, 4>(n, temp, temp_bytes, s);
  return path_of<8, 8>(n, temp, temp_bytes, s);
}

const char* last_stage_names() { return g_stage_names.c_str(); }

int sort_dispatch(mgs_key kt, mgs_val vt, const void* kin, void* kout, const void* vin, void* vout, size_t n,
                  void* temp, size_t temp_bytes, hipStream_t s, hipEvent_t* events, int num_events,
                  int* num_marks) {
  g_stage_names.clear();
  Marks m{events, num_events, 0, num_marks ? &g_stage_names : nullptr};
  int r = MGS_ERR_BAD_TYPE;
  switch (kt) {
    case MGS_KEY_U32: r = sort_by_val<MGS_KEY_U32>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s, m); break;
    case MGS_KEY_U64: r = sort_by_val<MGS_KEY_U64>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s, m); break;
    case MGS_KEY_I32: r = sort_by_val<MGS_KEY_I32>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s, m); break;
    case MGS_KEY_I64: r = sort_by_val<MGS_KEY_I64>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s, m); break;
    case MGS_KEY_F32: r = sort_by_val<MGS_KEY_F32>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s, m); break;
    case MGS_KEY_F64: r = sort_by_val<MGS_KEY_F64>(vt, kin, kout, vin, vout, n, temp, temp_bytes, s, m); break;
  }
  if (num_marks) *num_marks = m.used;
  return r;
}

}  // namespace mgs
