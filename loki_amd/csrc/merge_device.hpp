// Device helpers shared by the merge kernels: XCD-aware block remapping and
// the rotated float4 read of a parent row.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace loki_amd::kern {

/// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; remap
/// so XCD x works on the contiguous logical range [x*per(+rem), ...) and the
/// parent rows shared by neighbouring tiles stay inside one XCD's L2.  The
/// map is a bijection on [0, nblocks).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t kXcd = 8;
    const uint32_t per = nblocks / kXcd;
    const uint32_t rem = nblocks % kXcd;
    const uint32_t xcd = bid % kXcd;
    const uint32_t loc = bid / kXcd;
    return xcd < rem ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
}

/// Elements r..r+3 of the 8-element concatenation (a, b).
__device__ __forceinline__ float4 funnel4(const float4 a, const float4 b, uint32_t r) {
    float4 o;
    o.x = r == 0 ? a.x : (r == 1 ? a.y : (r == 2 ? a.z : a.w));
    o.y = r == 0 ? a.y : (r == 1 ? a.z : (r == 2 ? a.w : b.x));
    o.z = r == 0 ? a.z : (r == 1 ? a.w : (r == 2 ? b.x : b.y));
    o.w = r == 0 ? a.w : (r == 1 ? b.x : (r == 2 ? b.y : b.z));
    return o;
}

/// kernels.cpp:78-85: integer shift u32(shift + 0.5f), nbins -> 0.
__device__ __forceinline__ uint32_t shift_int(float s, uint32_t nbins) {
    const uint32_t k = static_cast<uint32_t>(s + 0.5F);
    return k == nbins ? 0U : k;
}

/// Output bins 4*i4 .. 4*i4+3 of one component row rotated right by S
/// (out[j] = row[(j - S) mod NB]); a rotated float4 straddles at most two
/// aligned float4 chunks.
template <int NB>
__device__ __forceinline__ float4 rotated4(const float4* __restrict__ row, uint32_t i4,
                                           uint32_t S) {
    constexpr uint32_t C4 = NB / 4;
    const uint32_t e0 = (4U * i4 + NB - S) & (NB - 1);
    const uint32_t c0 = e0 >> 2;
    const uint32_t c1 = (c0 + 1) & (C4 - 1);
    const float4 a    = row[c0];
    const float4 b    = row[c1];
    return funnel4(a, b, e0 & 3U);
}

}  // namespace loki_amd::kern
