// Merge path, Tanasic pivot, range swap and filler kernels for gfx950.
//
// Merge: Green, McColl & Bader, "GPU Merge Path" (2012).  Two partition
// kernels find, for every output tile boundary (diagonal), how many elements
// of A precede it (every 16th boundary over the whole inputs first, then the
// others inside the bracket of their coarse neighbours); each workgroup then
// stages its A and B slices in LDS, every
// thread co-ranks its own run of outputs inside the tile, merges it serially
// and the tile is written back with consecutive lanes on consecutive
// addresses.  Ties go to A (stable), as thrust::merge does at reference
// src/tanasic_sort.cuh:163-167.
#include <type_traits>

#include "common.hpp"
#include "internal.hpp"

namespace mgs {

namespace {

#ifndef MGS_MERGE_BLOCK
#define MGS_MERGE_BLOCK 256
#endif
#ifndef MGS_MERGE_IPT
#define MGS_MERGE_IPT 15
#endif
constexpr int kMergeBlock = MGS_MERGE_BLOCK;
constexpr int kMergeIpt = MGS_MERGE_IPT;  // odd: staggers LDS banks of the serial merges
constexpr int kMergeTile = kMergeBlock * kMergeIpt;
static_assert(kMergeIpt % 2 == 1 && kMergeTile >= 3840, "odd items per thread; tiles of at least 3840 keys");

// Tile boundaries per coarse split: the partition first co-ranks every
// kCoarse-th boundary over the whole inputs, then every boundary inside the
// bracket its two coarse neighbours give (the split is monotone in the
// diagonal for sorted inputs): about half the dependent global loads, and
// the fine searches of one bracket share cache lines.
constexpr uint64_t kCoarse = 16;

// Number of A elements among the first `diag` outputs of the stable merge,
// searched inside [lo_bound, hi_bound] (pass 0 and na for no bracket).
template <int KT, typename Bits>
__device__ __forceinline__ uint64_t corank(const Bits* a, uint64_t na, const Bits* b, uint64_t nb, uint64_t diag,
                                           uint64_t lo_bound = 0, uint64_t hi_bound = ~0ull) {
  uint64_t lo = diag > nb ? diag - nb : 0;
  uint64_t hi = diag < na ? diag : na;
  lo = std::max(lo, std::min(lo_bound, hi));
  hi = std::max(lo, std::min(hi, hi_bound));
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    // A[mid] is taken before B[diag-mid-1] iff enc(A[mid]) <= enc(B[diag-mid-1])
    if (Codec<KT>::enc(a[mid]) <= Codec<KT>::enc(b[diag - mid - 1]))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// coarse[q]: the split of tile boundary q kCoarse, for q <= num_coarse.
template <int KT>
__global__ void merge_coarse_kernel(const typename Codec<KT>::Bits* a, uint64_t na,
                                    const typename Codec<KT>::Bits* b, uint64_t nb, uint64_t* coarse,
                                    uint64_t num_coarse) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q > num_coarse) return;
  const uint64_t diag = std::min<uint64_t>(q * kCoarse * kMergeTile, na + nb);
  coarse[q] = corank<KT>(a, na, b, nb, diag);
}

// splits[i]: the split of tile boundary i, searched between the coarse
// splits around it.
template <int KT>
__global__ void merge_partition_kernel(const typename Codec<KT>::Bits* a, uint64_t na,
                                       const typename Codec<KT>::Bits* b, uint64_t nb, uint64_t* splits,
                                       uint64_t num_tiles, const uint64_t* __restrict__ coarse,
                                       uint64_t num_coarse) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > num_tiles) return;
  const uint64_t diag = std::min<uint64_t>(i * kMergeTile, na + nb);
  const uint64_t q = i / kCoarse;
  const uint64_t lo = coarse[q], hi = coarse[std::min(q + 1, num_coarse)];
  splits[i] = corank<KT>(a, na, b, nb, diag, lo, std::max(lo, hi));
}

template <int KT, typename V>
__global__ __launch_bounds__(kMergeBlock) void merge_tile_kernel(const typename Codec<KT>::Bits* __restrict__ a,
                                                                 const V* __restrict__ av, uint64_t na,
                                                                 const typename Codec<KT>::Bits* __restrict__ b,
                                                                 const V* __restrict__ bv, uint64_t nb,
                                                                 typename Codec<KT>::Bits* __restrict__ out,
                                                                 V* __restrict__ ov,
                                                                 const uint64_t* __restrict__ splits) {
  using Bits = typename Codec<KT>::Bits;
  constexpr bool HV = kHasVal<V>;
  using VS = typename std::conditional<HV, V, uint32_t>::type;
  __shared__ Bits sk[kMergeTile];
  __shared__ VS sv[HV ? kMergeTile : 1];

  const int tid = threadIdx.x;
  const uint64_t d0 = (uint64_t)blockIdx.x * kMergeTile;
  const uint64_t d1 = std::min<uint64_t>(d0 + kMergeTile, na + nb);
  // Splits are monotone for sorted inputs; the clamps only keep unsorted
  // (invalid) input from indexing outside A and B.
  const uint64_t a0 = min(splits[blockIdx.x], na);
  const uint64_t a1 = min(max(splits[blockIdx.x + 1], a0), na);
  const uint64_t b0 = d0 - a0;
  const int cnt = (int)(d1 - d0);
  const int la = (int)(a1 - a0);
  const int lb = cnt - la;

  // staging: every load of the thread is issued before the first LDS store
  {
    Bits lk[kMergeIpt];
    VS lv[kMergeIpt];
#pragma unroll
    for (int j = 0; j < kMergeIpt; ++j) {
      const int i = j * kMergeBlock + tid;
      const uint64_t jb = b0 + (uint64_t)(i - la);
      const bool from_a = i < la;
      const bool ok = i < cnt && (from_a || jb < nb);
      lk[j] = ok ? (from_a ? a[a0 + i] : b[jb]) : Bits(0);
      if constexpr (HV) lv[j] = ok ? (from_a ? av[a0 + i] : bv[jb]) : VS(0);
    }
#pragma unroll
    for (int j = 0; j < kMergeIpt; ++j) {
      const int i = j * kMergeBlock + tid;
      if (i < cnt) {
        sk[i] = lk[j];
        if constexpr (HV) sv[i] = lv[j];
      }
    }
  }
  __syncthreads();

  // Co-rank of this thread's first output inside the tile.
  const int diag = std::min(tid * kMergeIpt, cnt);
  int lo = diag > lb ? diag - lb : 0;
  int hi = diag < la ? diag : la;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (Codec<KT>::enc(sk[mid]) <= Codec<KT>::enc(sk[la + diag - mid - 1]))
      lo = mid + 1;
    else
      hi = mid;
  }
  int ia = lo, ib = la + (diag - lo);
  const int ea = la, eb = cnt;

  Bits rk[kMergeIpt];
  int src[kMergeIpt];
  const int mine = std::max(0, std::min(kMergeIpt, cnt - diag));
  Bits ka = ia < ea ? sk[ia] : Bits(0);
  Bits kb = ib < eb ? sk[ib] : Bits(0);
#pragma unroll
  for (int j = 0; j < kMergeIpt; ++j) {
    if (j < mine) {
      const bool take_a = (ib >= eb) || (ia < ea && Codec<KT>::enc(ka) <= Codec<KT>::enc(kb));
      if (take_a) {
        rk[j] = ka;
        src[j] = ia++;
        if (ia < ea) ka = sk[ia];
      } else {
        rk[j] = kb;
        src[j] = ib++;
        if (ib < eb) kb = sk[ib];
      }
    }
  }
  VS rv[kMergeIpt];
  if constexpr (HV) {
#pragma unroll
    for (int j = 0; j < kMergeIpt; ++j)
      if (j < mine) rv[j] = sv[src[j]];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kMergeIpt; ++j) {
    if (j < mine) {
      sk[diag + j] = rk[j];
      if constexpr (HV) sv[diag + j] = rv[j];
    }
  }
  __syncthreads();
  // write-out: every LDS read of the thread first, then the stores
#pragma unroll
  for (int j = 0; j < kMergeIpt; ++j) {
    const int i = j * kMergeBlock + tid;
    if (i < cnt) {
      rk[j] = sk[i];
      if constexpr (HV) rv[j] = sv[i];
    }
  }
#pragma unroll
  for (int j = 0; j < kMergeIpt; ++j) {
    const int i = j * kMergeBlock + tid;
    if (i < cnt) {
      out[d0 + i] = rk[j];
      if constexpr (HV) ov[d0 + i] = rv[j];
    }
  }
}

template <int KT, int VT>
int merge_typed(const void* ak, const void* av, size_t na, const void* bk, const void* bv, size_t nb, void* ok,
                void* ov, void* temp, size_t temp_bytes, hipStream_t s) {
  using Bits = typename Codec<KT>::Bits;
  using V = typename ValOf<VT>::T;
  const size_t total = na + nb;
  if (total == 0) return MGS_OK;
  const uint64_t tiles = (total + kMergeTile - 1) / kMergeTile;
  const uint64_t num_coarse = (tiles + kCoarse - 1) / kCoarse;
  if (temp == nullptr || temp_bytes < (tiles + 1 + num_coarse + 1) * sizeof(uint64_t)) return MGS_ERR_TEMP_TOO_SMALL;
  uint64_t* splits = static_cast<uint64_t*>(temp);
  uint64_t* coarse = splits + tiles + 1;
  const Bits* a = static_cast<const Bits*>(ak);
  const Bits* b = static_cast<const Bits*>(bk);
  merge_coarse_kernel<KT><<<(unsigned)((num_coarse + 1 + 255) / 256), 256, 0, s>>>(a, na, b, nb, coarse, num_coarse);
  merge_partition_kernel<KT><<<(unsigned)((tiles + 1 + 255) / 256), 256, 0, s>>>(a, na, b, nb, splits, tiles, coarse,
                                                                                   num_coarse);
  merge_tile_kernel<KT, V><<<(unsigned)tiles, kMergeBlock, 0, s>>>(
      a, static_cast<const V*>(av), na, b, static_cast<const V*>(bv), nb, static_cast<Bits*>(ok),
      static_cast<V*>(ov), splits);
  return check_launch();
}

// ---------------------------------------------------------------------------
// Pivot: 64-ary search by one wavefront.  Each round every lane evaluates the
// reference predicate at one candidate k (reference src/tanasic_sort.cuh:27-40:
// k is exchanged iff L[M-k] > R[k-1]); the predicate is true on a prefix of
// [0, M] for sorted partitions, so the first false lane brackets the answer
// and the interval shrinks ~64x per round (5 rounds for M = 2^29 instead of
// ~30 dependent loads).
// ---------------------------------------------------------------------------
struct PartPtrs {
  const void* p[8];
};

template <int KT>
__device__ __forceinline__ typename Codec<KT>::Bits virt_load(const PartPtrs& parts, uint64_t part_size, uint64_t i) {
  using Bits = typename Codec<KT>::Bits;
  const uint64_t q = i / part_size;
  return static_cast<const Bits*>(parts.p[q])[i - q * part_size];
}

template <int KT>
__global__ __launch_bounds__(64) void pivot_kernel(PartPtrs L, PartPtrs R, uint64_t part_size, uint64_t num_parts,
                                                   uint64_t* result) {
  const int lane = threadIdx.x;
  const uint64_t M = part_size * num_parts;
  uint64_t lo = 0, hi = M + 1;  // pred(lo) holds (pred(0) by definition), pred(hi) fails
  while (hi - lo > 1) {
    const uint64_t span = hi - lo - 1;  // unknown candidates lo+1 .. hi-1
    const uint64_t c = lo + 1 + (span / 64) * (uint64_t)lane + ((span % 64) * (uint64_t)lane) / 64;
    bool f = false;
    if (c <= M) f = Codec<KT>::enc(virt_load<KT>(L, part_size, M - c)) > Codec<KT>::enc(virt_load<KT>(R, part_size, c - 1));
    const uint64_t m = __ballot(f);
    const int t = (~m == 0) ? 64 : __builtin_ctzll(~m);  // first failing lane
    const uint64_t c_lo = __shfl(c, t > 0 ? t - 1 : 0, 64);
    const uint64_t c_hi = __shfl(c, t < 64 ? t : 63, 64);
    if (t > 0) lo = c_lo;
    if (t < 64) hi = c_hi;
  }
  if (lane == 0) {
    *result = lo;
    __threadfence_system();
  }
}

int pivot_typed_select(mgs_key kt, const PartPtrs& L, const PartPtrs& R, uint64_t part_size, uint64_t parts,
                       uint64_t* result, hipStream_t s) {
  switch (kt) {
    case MGS_KEY_U32: pivot_kernel<MGS_KEY_U32><<<1, 64, 0, s>>>(L, R, part_size, parts, result); break;
    case MGS_KEY_U64: pivot_kernel<MGS_KEY_U64><<<1, 64, 0, s>>>(L, R, part_size, parts, result); break;
    case MGS_KEY_I32: pivot_kernel<MGS_KEY_I32><<<1, 64, 0, s>>>(L, R, part_size, parts, result); break;
    case MGS_KEY_I64: pivot_kernel<MGS_KEY_I64><<<1, 64, 0, s>>>(L, R, part_size, parts, result); break;
    case MGS_KEY_F32: pivot_kernel<MGS_KEY_F32><<<1, 64, 0, s>>>(L, R, part_size, parts, result); break;
    case MGS_KEY_F64: pivot_kernel<MGS_KEY_F64><<<1, 64, 0, s>>>(L, R, part_size, parts, result); break;
    default: return MGS_ERR_BAD_TYPE;
  }
  return check_launch();
}

// ---------------------------------------------------------------------------
// Swap and fill.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void swap_kernel(T* __restrict__ a, T* __restrict__ b, size_t n) {
  const size_t st = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
    const T x = a[i];
    a[i] = b[i];
    b[i] = x;
  }
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, size_t n, T v) {
  const size_t st = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) p[i] = v;
}

// Swaps n elements of `bytes` each, using 16-byte words when both ranges
// allow it.
int swap_bytes(void* a, void* b, size_t n, int bytes, hipStream_t s) {
  if (n == 0) return MGS_OK;
  const size_t total = n * (size_t)bytes;
  const uintptr_t ua = reinterpret_cast<uintptr_t>(a), ub = reinterpret_cast<uintptr_t>(b);
  auto grid = [](size_t m) { return (unsigned)std::min<size_t>((m + 255) / 256, 8192); };
  if (ua % 16 == 0 && ub % 16 == 0 && total % 16 == 0) {
    const size_t m = total / 16;
    swap_kernel<uint4><<<grid(m), 256, 0, s>>>(static_cast<uint4*>(a), static_cast<uint4*>(b), m);
  } else if (bytes == 8) {
    swap_kernel<uint64_t><<<grid(n), 256, 0, s>>>(static_cast<uint64_t*>(a), static_cast<uint64_t*>(b), n);
  } else {
    swap_kernel<uint32_t><<<grid(n), 256, 0, s>>>(static_cast<uint32_t*>(a), static_cast<uint32_t*>(b), n);
  }
  return check_launch();
}

}  // namespace

size_t merge_temp_bytes(mgs_key kt, mgs_val vt, size_t na, size_t nb) {
  if (key_bytes(kt) < 0 || val_bytes(vt) < 0) return 0;
  const size_t tiles = (na + nb + kMergeTile - 1) / kMergeTile;
  return (tiles + 1 + (tiles + kCoarse - 1) / kCoarse + 1) * sizeof(uint64_t);
}

int merge_dispatch(mgs_key kt, mgs_val vt, const void* ak, const void* av, size_t na, const void* bk, const void* bv,
                   size_t nb, void* ok, void* ov, void* temp, size_t temp_bytes, hipStream_t s) {
#define MGS_MERGE_V(KT)                                                                                      \
  switch (vt) {                                                                                              \
    case MGS_VAL_NONE: return merge_typed<KT, MGS_VAL_NONE>(ak, av, na, bk, bv, nb, ok, ov, temp, temp_bytes, s); \
    case MGS_VAL_U32: return merge_typed<KT, MGS_VAL_U32>(ak, av, na, bk, bv, nb, ok, ov, temp, temp_bytes, s);   \
    case MGS_VAL_U64: return merge_typed<KT, MGS_VAL_U64>(ak, av, na, bk, bv, nb, ok, ov, temp, temp_bytes, s);   \
  }                                                                                                          \
  return MGS_ERR_BAD_TYPE;
  switch (kt) {
    case MGS_KEY_U32: MGS_MERGE_V(MGS_KEY_U32)
    case MGS_KEY_U64: MGS_MERGE_V(MGS_KEY_U64)
    case MGS_KEY_I32: MGS_MERGE_V(MGS_KEY_I32)
    case MGS_KEY_I64: MGS_MERGE_V(MGS_KEY_I64)
    case MGS_KEY_F32: MGS_MERGE_V(MGS_KEY_F32)
    case MGS_KEY_F64: MGS_MERGE_V(MGS_KEY_F64)
  }
#undef MGS_MERGE_V
  return MGS_ERR_BAD_TYPE;
}

int pivot_dispatch(mgs_key kt, const void* const* left, const void* const* right, size_t part_size, size_t num_parts,
                   uint64_t* result, hipStream_t s) {
  if (num_parts == 0 || num_parts > 8 || left == nullptr || right == nullptr || result == nullptr)
    return MGS_ERR_INVALID;
  PartPtrs L{}, R{};
  for (size_t i = 0; i < num_parts; ++i) {
    if (part_size > 0 && (left[i] == nullptr || right[i] == nullptr)) return MGS_ERR_INVALID;
    L.p[i] = left[i];
    R.p[i] = right[i];
  }
  return pivot_typed_select(kt, L, R, part_size, num_parts, result, s);
}

int swap_dispatch(mgs_key kt, mgs_val vt, void* ak, void* av, void* bk, void* bv, size_t n, hipStream_t s) {
  const int kb = key_bytes(kt), vb = val_bytes(vt);
  if (kb < 0 || vb < 0) return MGS_ERR_BAD_TYPE;
  if (n == 0) return MGS_OK;
  if (ak == nullptr || bk == nullptr || (vb > 0 && (av == nullptr || bv == nullptr))) return MGS_ERR_INVALID;
  int st = swap_bytes(ak, bk, n, kb, s);
  if (st != MGS_OK || vb == 0) return st;
  return swap_bytes(av, bv, n, vb, s);
}

int fill_max_dispatch(mgs_key kt, void* keys, size_t n, hipStream_t s) {
  if (n == 0) return MGS_OK;
  if (keys == nullptr) return MGS_ERR_INVALID;
  const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  switch (kt) {
    case MGS_KEY_U32: fill_kernel<uint32_t><<<grid, 256, 0, s>>>((uint32_t*)keys, n, Codec<MGS_KEY_U32>::dec(~0u)); break;
    case MGS_KEY_I32: fill_kernel<uint32_t><<<grid, 256, 0, s>>>((uint32_t*)keys, n, Codec<MGS_KEY_I32>::dec(~0u)); break;
    case MGS_KEY_F32: fill_kernel<uint32_t><<<grid, 256, 0, s>>>((uint32_t*)keys, n, Codec<MGS_KEY_F32>::dec(~0u)); break;
    case MGS_KEY_U64: fill_kernel<uint64_t><<<grid, 256, 0, s>>>((uint64_t*)keys, n, Codec<MGS_KEY_U64>::dec(~0ull)); break;
    case MGS_KEY_I64: fill_kernel<uint64_t><<<grid, 256, 0, s>>>((uint64_t*)keys, n, Codec<MGS_KEY_I64>::dec(~0ull)); break;
    case MGS_KEY_F64: fill_kernel<uint64_t><<<grid, 256, 0, s>>>((uint64_t*)keys, n, Codec<MGS_KEY_F64>::dec(~0ull)); break;
    default: return MGS_ERR_BAD_TYPE;
  }
  return check_launch();
}

}  // namespace mgs
