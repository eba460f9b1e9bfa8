// Fold storage precision.  Folds of levels 0..L-2 may be stored as FP16 or
// BF16; every add is done in FP32 and each stored value is rounded once
// (round to nearest even).  Level l is stored at the exact power-of-two scale
// 2^-(k0 + l): a merge is then (tail + head) * 0.5, which keeps the stored
// magnitudes of order one (no FP16 overflow of the growing sums) and is
// exact, so unscaling the last level by 2^(k0 + L - 2) restores the true
// values without rounding.
//
// Block floating point (BFP16) stores each component row (the nbins e or v
// values of one segment and coordinate) as nbins int16 mantissas m and one
// shared exponent E, value = m * 2^E, with E = max(ex(M) - 14, -126) for the
// row maximum M = max |x| (ex: unbiased exponent field, -127 for zero or
// subnormal M) and m = rint(x * 2^-E), ties to even, clamped to +-32767: 15
// significant bits relative to the row's largest value, where FP16 keeps 11
// relative to each value.  Levels 1..L-2 are stored this way, level 0 (the
// brute fold) in FP32, with no level scaling.  A BFP16 level of n elements
// is the n int16 mantissas followed by its exponent plane: one 16-bit word
// per (segment, coordinate) row pair, low byte the e row's E, high byte the
// v row's (n / nbins bytes).  Inputs must be finite.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace loki_amd {

enum class FoldStorage : int { kF32 = 0, kF16 = 1, kBF16 = 2, kBFP16 = 3 };

/// Bytes per stored element (BFP16: the mantissa; its exponent plane is
/// counted by level_bytes).
inline uint32_t storage_bytes(FoldStorage s) { return s == FoldStorage::kF32 ? 4U : 2U; }

/// Bytes of a stored level of `elems` elements in rows of nbins.
inline uint64_t level_bytes(FoldStorage s, uint64_t elems, uint64_t nbins) {
    return s == FoldStorage::kBFP16 ? 2 * elems + elems / nbins : elems * storage_bytes(s);
}

namespace kern {

/// Raw chunk of 4 consecutive stored elements.  BFP16 mantissas are int16_t.
template <typename T> struct Chunk4;
template <> struct Chunk4<float> { using type = float4; };
template <> struct Chunk4<_Float16> { using type = uint2; };
template <> struct Chunk4<__bf16> { using type = uint2; };
template <> struct Chunk4<int16_t> { using type = uint2; };
template <typename T> using chunk4_t = typename Chunk4<T>::type;

/// Stored type is block floating point (int16 mantissas + row exponents).
template <typename T> constexpr bool kBlockFP = std::is_same_v<T, int16_t>;
/// Stored type is FP16 or BF16, kept at the level's power-of-two scale.
template <typename T>
constexpr bool kScaledHalf = std::is_same_v<T, _Float16> || std::is_same_v<T, __bf16>;

/// 2^e as FP32, e in [-126, 127].
__device__ __forceinline__ float pow2f(int e) {
    return __uint_as_float(static_cast<uint32_t>(e + 127) << 23);
}

/// Shared exponent of a BFP16 row whose largest magnitude is mx.
__device__ __forceinline__ int bfp_exponent(float mx) {
    const int ex = static_cast<int>((__float_as_uint(mx) >> 23) & 0xFFU) - 127;
    return ex - 14 < -126 ? -126 : ex - 14;
}

__device__ __forceinline__ float i16_lo(uint32_t w) {
    return static_cast<float>(static_cast<int>(static_cast<int16_t>(w & 0xFFFFU)));
}
__device__ __forceinline__ float i16_hi(uint32_t w) {
    return static_cast<float>(static_cast<int>(w) >> 16);
}

/// 4 BFP16 mantissas -> FP32 values m * s (s = 2^E; exact).
__device__ __forceinline__ float4 widen4_bfp(const uint2 c, float s) {
    return make_float4(i16_lo(c.x) * s, i16_hi(c.x) * s, i16_lo(c.y) * s, i16_hi(c.y) * s);
}

/// rint(x * inv) clamped to +-32767 as a 16-bit pattern (inv = 2^-E).
__device__ __forceinline__ uint32_t bfp_mant(float x, float inv) {
    const float m = __builtin_fminf(__builtin_fmaxf(__builtin_rintf(x * inv), -32767.0F), 32767.0F);
    return static_cast<uint32_t>(static_cast<int>(m)) & 0xFFFFU;
}

/// 4 FP32 values -> BFP16 mantissas at inv = 2^-E.
__device__ __forceinline__ uint2 narrow4_bfp(const float4 v, float inv) {
    return make_uint2(bfp_mant(v.x, inv) | (bfp_mant(v.y, inv) << 16),
                      bfp_mant(v.z, inv) | (bfp_mant(v.w, inv) << 16));
}

/// FP32 value as stored in BFP16 at exponent E (the decoded rounding).
__device__ __forceinline__ float round_bfp(float x, float inv, float s) {
    const float m = __builtin_rintf(x * inv);
    return __builtin_fminf(__builtin_fmaxf(m, -32767.0F), 32767.0F) * s;
}

/// Largest |x| over aligned groups of G consecutive lanes (G a power of two
/// <= 64); every lane of a group must be active.
template <int G>
__device__ __forceinline__ float group_absmax(float x) {
    x = __builtin_fabsf(x);
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
        x = __builtin_fmaxf(x, __shfl_xor(x, o));
    }
    return x;
}

/// Streaming (non-temporal) store of one chunk: fold levels are far larger
/// than the caches and are next read by another launch, so their writes should
/// not evict the parent rows that the current launch is still re-reading.
__device__ __forceinline__ void store_stream(float4* p, const float4 v) {
    using v4 = float __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store(v4{v.x, v.y, v.z, v.w}, reinterpret_cast<v4*>(p));
}
__device__ __forceinline__ void store_stream(uint2* p, const uint2 v) {
    using u2 = unsigned int __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(u2{v.x, v.y}, reinterpret_cast<u2*>(p));
}

__device__ __forceinline__ float f16_lo(uint32_t w) {
    return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w & 0xFFFFU)));
}
__device__ __forceinline__ float f16_hi(uint32_t w) {
    return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w >> 16)));
}
__device__ __forceinline__ uint32_t f16_bits(float x) {
    return static_cast<uint32_t>(__builtin_bit_cast(uint16_t, static_cast<_Float16>(x)));
}
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000U); }
__device__ __forceinline__ uint32_t bf16_bits(float x) {
    return static_cast<uint32_t>(__builtin_bit_cast(uint16_t, static_cast<__bf16>(x)));
}

/// 4 stored elements -> FP32 (exact).
template <typename T> __device__ __forceinline__ float4 widen4(const chunk4_t<T> c);
template <> __device__ __forceinline__ float4 widen4<float>(const float4 c) { return c; }
template <> __device__ __forceinline__ float4 widen4<_Float16>(const uint2 c) {
    return make_float4(f16_lo(c.x), f16_hi(c.x), f16_lo(c.y), f16_hi(c.y));
}
template <> __device__ __forceinline__ float4 widen4<__bf16>(const uint2 c) {
    return make_float4(bf16_lo(c.x), bf16_hi(c.x), bf16_lo(c.y), bf16_hi(c.y));
}

/// FP32 -> 4 stored elements (round to nearest even).
template <typename T> __device__ __forceinline__ chunk4_t<T> narrow4(const float4 v);
template <> __device__ __forceinline__ float4 narrow4<float>(const float4 v) { return v; }
template <> __device__ __forceinline__ uint2 narrow4<_Float16>(const float4 v) {
    return make_uint2(f16_bits(v.x) | (f16_bits(v.y) << 16), f16_bits(v.z) | (f16_bits(v.w) << 16));
}
template <> __device__ __forceinline__ uint2 narrow4<__bf16>(const float4 v) {
    return make_uint2(bf16_bits(v.x) | (bf16_bits(v.y) << 16),
                      bf16_bits(v.z) | (bf16_bits(v.w) << 16));
}

/// One stored element <-> FP32.
template <typename T> __device__ __forceinline__ float widen1(T x) {
    return static_cast<float>(x);
}
template <typename T> __device__ __forceinline__ T narrow1(float x) { return static_cast<T>(x); }

/// Elements r..r+3 of the 8 16-bit elements (a, b) (little-endian halves):
/// one select level on 32-bit words, then a 0- or 16-bit funnel shift.
__device__ __forceinline__ uint2 funnel4_16(const uint2 a, const uint2 b, uint32_t r) {
    const bool k      = (r & 2U) != 0;
    const uint32_t sh = (r & 1U) * 16U;
    const uint32_t w0 = k ? a.y : a.x;
    const uint32_t w1 = k ? b.x : a.y;
    const uint32_t w2 = k ? b.y : b.x;
    return make_uint2(__builtin_amdgcn_alignbit(w1, w0, sh), __builtin_amdgcn_alignbit(w2, w1, sh));
}

}  // namespace kern
}  // namespace loki_amd
