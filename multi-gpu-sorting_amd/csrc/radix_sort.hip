// Onesweep LSD radix sort kernels and their launch sequence (see
// radix_sort.hpp for the algorithm outline).
#include <type_traits>

#include "internal.hpp"
#include "radix_device.hpp"

namespace mgs {

// ---------------------------------------------------------------------------
// init: zero histograms and tile counters.
// ---------------------------------------------------------------------------
__global__ void init_kernel(uint32_t* hist, uint32_t* tile_counter, int passes) {
  for (int i = threadIdx.x; i < passes * kRadix; i += blockDim.x) hist[i] = 0;
  if (threadIdx.x < kMaxPasses) tile_counter[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// hist: 256-bin histogram of every byte of the encoded keys in one read.
// Also clears the look-back descriptor array (desc_words 32-bit words).
// ---------------------------------------------------------------------------
template <int KT, int BLOCK, bool VEC>
__global__ __launch_bounds__(BLOCK) void hist_kernel(const typename Codec<KT>::Bits* __restrict__ keys, uint32_t n,
                                                     uint32_t* __restrict__ hist, uint32_t* __restrict__ desc,
                                                     size_t desc_words) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int P = (int)sizeof(Bits);
  __shared__ uint32_t h[P][kRadix];
  const int tid = threadIdx.x;
  for (int i = tid; i < P * kRadix; i += BLOCK) (&h[0][0])[i] = 0;

  const size_t gtid = (size_t)blockIdx.x * BLOCK + tid;
  const size_t gstride = (size_t)gridDim.x * BLOCK;
  {  // clear descriptors with 16-byte stores
    uint4* d4 = reinterpret_cast<uint4*>(desc);
    const size_t n4 = desc_words / 4;
    for (size_t i = gtid; i < n4; i += gstride) d4[i] = make_uint4(0, 0, 0, 0);
    for (size_t i = n4 * 4 + gtid; i < desc_words; i += gstride) desc[i] = 0;
  }
  __syncthreads();

  auto count = [&](Bits k) {
    const Bits e = Codec<KT>::enc(k);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      uint32_t d;
      if constexpr (sizeof(Bits) == 8)
        d = (uint32_t)(e >> (8 * p)) & 0xFFu;
      else
        d = __builtin_amdgcn_ubfe((uint32_t)e, 8u * p, 8u);
      atomicAdd(&h[p][d], 1u);
    }
  };

  if constexpr (VEC) {
    constexpr int V = 16 / (int)sizeof(Bits);
    using Vec = typename std::conditional<sizeof(Bits) == 4, uint4, ulonglong2>::type;
    const Vec* kv = reinterpret_cast<const Vec*>(keys);
    const size_t nv = n / V;
    for (size_t i = gtid; i < nv; i += gstride) {
      Vec v = kv[i];
      if constexpr (sizeof(Bits) == 4) {
        count(v.x); count(v.y); count(v.z); count(v.w);
      } else {
        count(v.x); count(v.y);
      }
    }
    for (size_t i = nv * V + gtid; i < n; i += gstride) count(keys[i]);
  } else {
    for (size_t i = gtid; i < n; i += gstride) count(keys[i]);
  }
  __syncthreads();
  for (int i = tid; i < P * kRadix; i += BLOCK) {
    const uint32_t c = (&h[0][0])[i];
    if (c) atomicAdd(&hist[i], c);
  }
}

// ---------------------------------------------------------------------------
// plan: digit bases, trivial-pass detection, ping-pong assignment.
// Launched with exactly 256 threads.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void plan_kernel(const uint32_t* __restrict__ hist, uint32_t n, int passes,
                                                   int aliased, PassPlan* __restrict__ plan,
                                                   uint32_t* __restrict__ digit_base) {
  __shared__ int trivial[kMaxPasses];
  __shared__ uint32_t sums[4];
  const int tid = threadIdx.x;
  for (int p = 0; p < passes; ++p) {
    const uint32_t c = hist[p * kRadix + tid];
    const int t = __syncthreads_or(c == n);
    const uint32_t excl = block_excl_scan_256(c, sums, tid);
    digit_base[p * kRadix + tid] = excl;
    if (tid == 0) trivial[p] = t;
  }
  __syncthreads();
  if (tid == 0) {
    int na = 0;
    for (int p = 0; p < kMaxPasses; ++p) plan->k_of_pass[p] = -1;
    for (int p = 0; p < passes; ++p)
      if (!trivial[p]) plan->k_of_pass[p] = na++;
    plan->num_active = na;
    for (int k = 0; k < na; ++k) {
      int d;
      if (aliased)
        d = (k % 2 == 0) ? kBufAlt : kBufOut;
      else
        d = ((na - 1 - k) % 2 == 0) ? kBufOut : kBufAlt;
      plan->dst[k] = d;
      plan->src[k] = (k == 0) ? kBufIn : plan->dst[k - 1];
    }
    if (na == 0)
      plan->final_copy_src = kBufIn;
    else
      plan->final_copy_src = (plan->dst[na - 1] == kBufOut) ? -1 : plan->dst[na - 1];
  }
}

template <typename T>
__device__ __forceinline__ const T* pick_src(int id, const T* in, const T* alt, const T* out) {
  return id == kBufIn ? in : (id == kBufAlt ? alt : out);
}
template <typename T>
__device__ __forceinline__ T* pick_dst(int id, T* alt, T* out) {
  return id == kBufAlt ? alt : out;
}

// ---------------------------------------------------------------------------
// One onesweep pass over byte `pass` of the whole array.
// ---------------------------------------------------------------------------
template <int KT, typename V, int BLOCK, int IPT, typename D>
__global__ __launch_bounds__(BLOCK) void onesweep_kernel(
    const typename Codec<KT>::Bits* k_in, typename Codec<KT>::Bits* k_alt, typename Codec<KT>::Bits* k_out,
    const V* v_in, V* v_alt, V* v_out, uint32_t n, int pass, const PassPlan* __restrict__ plan,
    const uint32_t* __restrict__ digit_base, uint32_t* __restrict__ tile_counter, D* __restrict__ desc) {
  using Bits = typename Codec<KT>::Bits;
  const int ka = plan->k_of_pass[pass];
  if (ka < 0) return;
  const int sb = plan->src[ka], db = plan->dst[ka];
  const Bits* src = pick_src(sb, k_in, (const Bits*)k_alt, (const Bits*)k_out);
  Bits* dst = pick_dst(db, k_alt, k_out);
  const V* vsrc = nullptr;
  V* vdst = nullptr;
  if constexpr (kHasVal<V>) {
    vsrc = pick_src(sb, v_in, (const V*)v_alt, (const V*)v_out);
    vdst = pick_dst(db, v_alt, v_out);
  }
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(&tile_counter[pass], 1u);
  __syncthreads();
  onesweep_tile<KT, V, BLOCK, IPT, D>(src, dst, vsrc, vdst, n, s_tile, desc, digit_base + pass * kRadix, pass * 8,
                                      ka & 1);
}

// ---------------------------------------------------------------------------
// fixup: copy the result into the output buffers when the plan says so.
// ---------------------------------------------------------------------------
template <typename Bits, typename V>
__global__ void fixup_kernel(const Bits* k_in, const Bits* k_alt, Bits* k_out, const V* v_in, const V* v_alt, V* v_out,
                             uint32_t n, const PassPlan* __restrict__ plan) {
  const int s = plan->final_copy_src;
  if (s < 0) return;
  const Bits* ks = s == kBufIn ? k_in : k_alt;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  if (ks != k_out)
    for (size_t i = g; i < n; i += st) k_out[i] = ks[i];
  if constexpr (kHasVal<V>) {
    const V* vs = s == kBufIn ? v_in : v_alt;
    if (vs != v_out)
      for (size_t i = g; i < n; i += st) v_out[i] = vs[i];
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

struct SortLayout {
  size_t alt_k, alt_v, hist, digit_base, tile_counter, plan, desc, total;
  uint32_t num_tiles;
  bool wide_desc;
};

template <int KB, int VB>
SortLayout sort_layout(size_t n) {
  using Cfg = TileCfg<KB, VB>;
  SortLayout L{};
  size_t o = 0;
  L.alt_k = o;
  o = align_up(o + n * KB);
  L.alt_v = o;
  o = align_up(o + n * VB);
  L.hist = o;
  o = align_up(o + kMaxPasses * kRadix * 4);
  L.digit_base = o;
  o = align_up(o + kMaxPasses * kRadix * 4);
  L.tile_counter = o;
  o = align_up(o + kMaxPasses * 4);
  L.plan = o;
  o = align_up(o + sizeof(PassPlan));
  L.num_tiles = (uint32_t)((n + Cfg::kTile - 1) / Cfg::kTile);
  L.wide_desc = n > (size_t(1) << 30);
  L.desc = o;
  o = align_up(o + (size_t)L.num_tiles * kRadix * (L.wide_desc ? 8 : 4));
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
  const VP* vi = static_cast<const VP*>(vin);
  VP* vo = static_cast<VP*>(vout);

  if (n <= (size_t)SCfg::kTile) {
    small_sort_kernel<KT, V, SCfg::kBlock, SCfg::kIpt>
        <<<1, SCfg::kBlock, 0, s>>>(ki, ko, (const V*)vi, (V*)vo, (uint32_t)n);
    return check_launch();
  }

  const SortLayout L = sort_layout<KB, VB>(n);
  if (temp == nullptr || temp_bytes < L.total) return MGS_ERR_TEMP_TOO_SMALL;
  char* t = static_cast<char*>(temp);
  Bits* ka = reinterpret_cast<Bits*>(t + L.alt_k);
  VP* va = reinterpret_cast<VP*>(t + L.alt_v);
  uint32_t* hist = reinterpret_cast<uint32_t*>(t + L.hist);
  uint32_t* dbase = reinterpret_cast<uint32_t*>(t + L.digit_base);
  uint32_t* tcount = reinterpret_cast<uint32_t*>(t + L.tile_counter);
  PassPlan* plan = reinterpret_cast<PassPlan*>(t + L.plan);
  void* desc = t + L.desc;
  const size_t desc_words = (size_t)L.num_tiles * kRadix * (L.wide_desc ? 2 : 1);
  const int aliased = (kin == kout) || (kHasVal<V> && vin == vout);

  init_kernel<<<1, 256, 0, s>>>(hist, tcount, P);
  {
    constexpr int HB = 256;
    const size_t work = (n + 16 * HB - 1) / (16 * HB);
    const int grid = (int)std::min<size_t>(std::max<size_t>(work, 1), 2048);
    const bool vec = (reinterpret_cast<uintptr_t>(kin) % 16) == 0;
    if (vec)
      hist_kernel<KT, HB, true><<<grid, HB, 0, s>>>(ki, (uint32_t)n, hist, (uint32_t*)desc, desc_words);
    else
      hist_kernel<KT, HB, false><<<grid, HB, 0, s>>>(ki, (uint32_t)n, hist, (uint32_t*)desc, desc_words);
  }
  plan_kernel<<<1, 256, 0, s>>>(hist, (uint32_t)n, P, aliased, plan, dbase);
  for (int p = 0; p < P; ++p) {
    if (L.wide_desc)
      onesweep_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, uint64_t><<<L.num_tiles, Cfg::kBlock, 0, s>>>(
          ki, ka, ko, (const V*)vi, (V*)va, (V*)vo, (uint32_t)n, p, plan, dbase, tcount, (uint64_t*)desc);
    else
      onesweep_kernel<KT, V, Cfg::kBlock, Cfg::kIpt, uint32_t><<<L.num_tiles, Cfg::kBlock, 0, s>>>(
          ki, ka, ko, (const V*)vi, (V*)va, (V*)vo, (uint32_t)n, p, plan, dbase, tcount, (uint32_t*)desc);
  }
  {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    fixup_kernel<Bits, V><<<grid, 256, 0, s>>>(ki, ka, ko, (const V*)vi, (const V*)va, (V*)vo, (uint32_t)n, plan);
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
