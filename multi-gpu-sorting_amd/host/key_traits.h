// Host-side mapping from C++ key types to the libmgsort key enum, and the
// encoded-order comparison every CPU path uses to check results.
#pragma once

#include <cstdint>
#include <cstring>
#include <type_traits>

#include "mgsort.h"

namespace mgsort {

template <typename T> struct KeyOf;
template <> struct KeyOf<int32_t> { static constexpr mgs_key value = MGS_KEY_I32; };
template <> struct KeyOf<int64_t> { static constexpr mgs_key value = MGS_KEY_I64; };
template <> struct KeyOf<long long> { static constexpr mgs_key value = MGS_KEY_I64; };
template <> struct KeyOf<uint32_t> { static constexpr mgs_key value = MGS_KEY_U32; };
template <> struct KeyOf<uint64_t> { static constexpr mgs_key value = MGS_KEY_U64; };
template <> struct KeyOf<float> { static constexpr mgs_key value = MGS_KEY_F32; };
template <> struct KeyOf<double> { static constexpr mgs_key value = MGS_KEY_F64; };

// Unsigned image whose ascending order is the library's key order
// (-0.0 folded onto +0.0).
template <typename T>
inline auto Encode(T x) {
  if constexpr (std::is_floating_point_v<T>) {
    using U = std::conditional_t<sizeof(T) == 4, uint32_t, uint64_t>;
    U u;
    std::memcpy(&u, &x, sizeof(U));
    constexpr U sign = U(1) << (sizeof(U) * 8 - 1);
    if (u == sign) u = 0;
    return (u & sign) ? U(~u) : U(u | sign);
  } else if constexpr (std::is_signed_v<T>) {
    using U = std::make_unsigned_t<T>;
    return U(U(x) ^ (U(1) << (sizeof(U) * 8 - 1)));
  } else {
    return x;
  }
}

template <typename T>
struct EncodedLess {
  bool operator()(const T& a, const T& b) const { return Encode(a) < Encode(b); }
};

}  // namespace mgsort
