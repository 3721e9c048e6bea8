// Device helpers shared by the gfx950 sorting kernels: key codecs, 64-lane
// wavefront primitives and agent-scope atomics for decoupled look-back.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mgsort.h"

namespace mgs {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kRadixBits = 8;      // one byte per LSD pass
constexpr int kRadix = 1 << kRadixBits;

// ---------------------------------------------------------------------------
// Key codecs.  Keys travel through the kernels as their raw bits (uint32_t or
// uint64_t); enc() maps the bits onto an unsigned image whose ascending order
// is the sort order, dec() is its inverse (used only for the filler value).
// Floating point: -0.0 is folded onto +0.0 before the sign transform, so both
// zeros share one image and keep their input order under a stable sort.
// ---------------------------------------------------------------------------
template <int KT> struct Codec;

template <> struct Codec<MGS_KEY_U32> {
  using Bits = uint32_t;
  __host__ __device__ static inline Bits enc(Bits x) { return x; }
  __host__ __device__ static inline Bits dec(Bits u) { return u; }
};
template <> struct Codec<MGS_KEY_U64> {
  using Bits = uint64_t;
  __host__ __device__ static inline Bits enc(Bits x) { return x; }
  __host__ __device__ static inline Bits dec(Bits u) { return u; }
};
template <> struct Codec<MGS_KEY_I32> {
  using Bits = uint32_t;
  __host__ __device__ static inline Bits enc(Bits x) { return x ^ 0x80000000u; }
  __host__ __device__ static inline Bits dec(Bits u) { return u ^ 0x80000000u; }
};
template <> struct Codec<MGS_KEY_I64> {
  using Bits = uint64_t;
  __host__ __device__ static inline Bits enc(Bits x) { return x ^ 0x8000000000000000ull; }
  __host__ __device__ static inline Bits dec(Bits u) { return u ^ 0x8000000000000000ull; }
};
template <> struct Codec<MGS_KEY_F32> {
  using Bits = uint32_t;
  __host__ __device__ static inline Bits enc(Bits x) {
    x = (x == 0x80000000u) ? 0u : x;                           // -0.0 -> +0.0
    Bits m = (Bits)((int32_t)x >> 31) | 0x80000000u;           // neg: ~x, pos: x|sign
    return x ^ m;
  }
  __host__ __device__ static inline Bits dec(Bits u) {
    Bits m = (u & 0x80000000u) ? 0x80000000u : 0xFFFFFFFFu;
    return u ^ m;
  }
};
template <> struct Codec<MGS_KEY_F64> {
  using Bits = uint64_t;
  __host__ __device__ static inline Bits enc(Bits x) {
    x = (x == 0x8000000000000000ull) ? 0ull : x;
    Bits m = (Bits)((int64_t)x >> 63) | 0x8000000000000000ull;
    return x ^ m;
  }
  __host__ __device__ static inline Bits dec(Bits u) {
    Bits m = (u & 0x8000000000000000ull) ? 0x8000000000000000ull : ~0ull;
    return u ^ m;
  }
};

// Byte `pass` of the encoded key (pass 0 = least significant byte).
template <int KT>
__device__ __forceinline__ uint32_t digit_of(typename Codec<KT>::Bits raw, int shift) {
  typename Codec<KT>::Bits e = Codec<KT>::enc(raw);
  if constexpr (sizeof(e) == 8) {
    uint32_t half = (shift >= 32) ? (uint32_t)(e >> 32) : (uint32_t)e;
    return __builtin_amdgcn_ubfe(half, (uint32_t)(shift & 31), 8u);
  } else {
    return __builtin_amdgcn_ubfe((uint32_t)e, (uint32_t)shift, 8u);
  }
}

// Payload type helper: NoVal marks keys-only sorts.
struct NoVal {};
template <int VT> struct ValOf { using T = NoVal; };
template <> struct ValOf<MGS_VAL_U32> { using T = uint32_t; };
template <> struct ValOf<MGS_VAL_U64> { using T = uint64_t; };
template <typename V> constexpr bool kHasVal = !__is_same(V, NoVal);
template <typename V> constexpr int kValBytes = kHasVal<V> ? (int)sizeof(V) : 0;

// ---------------------------------------------------------------------------
// Wavefront helpers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// popcount(mask & lanes-below-me)
__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t readfirstlane(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Inclusive prefix sum over the 64 lanes of a wavefront, in DPP moves (no
// LDS): row_shr 1, 2, 4, 8 scan each row of 16 lanes (a source lane outside
// the row reads as 0), then row_bcast:15 adds row 0's total to row 1 and row
// 2's to row 3, and row_bcast:31 adds the first two rows' total to rows 2
// and 3 (row masks 0xA and 0xC leave the other rows' values at 0).  Every
// lane of the wavefront must be active.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

__device__ __forceinline__ uint32_t wave_broadcast(uint32_t v, int src_lane) { return __shfl(v, src_lane, kWave); }

// ---------------------------------------------------------------------------
// Agent-scope (cross-XCD coherent) relaxed atomics for the look-back chain.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T ld_agent(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_agent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Index of the XCD (one of the accelerator dies, each with its own L2) the
// wavefront runs on.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcc_id() {
  uint32_t v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return v & 0xFu;
}

// Workgroup-scope relaxed atomics on LDS tables that lanes of one wavefront
// update and re-read within an instruction stream (see rank_wave()).
template <typename T>
__device__ __forceinline__ T ld_wg(T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void st_wg(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void or_wg(T* p, T v) {
  __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

}  // namespace mgs
