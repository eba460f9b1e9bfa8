// Device helpers of the boxcar S/N (score.cpp:95-172, utils.cpp:22-31).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.hpp"

namespace loki_amd::kern {

/// max_j (hi[j] - ps[j]) over j < NB.  fmaxf pairs lower to v_max3_f32.  For
/// non-NaN differences this equals the reference's running
/// std::max(max_diff, d) (utils.cpp:22-31); NaN differences are skipped by
/// both.  The only conceivable divergence is which zero is kept when +0.0
/// and -0.0 tie, which cannot change the S/N value.
template <int NB>
__device__ __forceinline__ float window_max(const float* __restrict__ hi, const float* ps) {
    // two running maxima, each taking two differences per step (one v_max3_f32)
    float ma = -FLT_MAX, mb = -FLT_MAX;
    int j = 0;
#pragma unroll
    for (; j + 3 < NB; j += 4) {
        ma = __builtin_fmaxf(__builtin_fmaxf(ma, hi[j] - ps[j]), hi[j + 1] - ps[j + 1]);
        mb = __builtin_fmaxf(__builtin_fmaxf(mb, hi[j + 2] - ps[j + 2]), hi[j + 3] - ps[j + 3]);
    }
#pragma unroll
    for (; j < NB; ++j) {
        ma = __builtin_fmaxf(ma, hi[j] - ps[j]);
    }
    return __builtin_fmaxf(ma, mb);
}

/// Profile normalisation p = e / sqrt(v).
/// EXACT: the CPU reference's e * (1.0f / sqrtf(v)) with correctly rounded
/// sqrt and division (score.cpp:139-141) -- bit-identical to the oracle.
/// Otherwise one hardware reciprocal square root (v_rsq_f32, ~1 ulp), the
/// analogue of the reference CUDA path's fast-math e / sqrtf(v)
/// (score_cuda.cu:82, built with -use_fast_math).
template <bool EXACT>
__device__ __forceinline__ float normalise(float e, float v) {
    if constexpr (EXACT) {
        return e * (1.0F / sqrtf(v));
    } else {
        return e * __builtin_amdgcn_rsqf(v);
    }
}

/// Appends (index, snr) of every active lane whose `pass` is set to a
/// compacted candidate list (score_cuda.cu:130-168 kPerWidthAndFilter /
/// kMaxAndFilter): one atomic add per wave reserves the slots, lanes fill them
/// in lane order.  Slots at or beyond fo.capacity are counted but not written.
/// Must be reached by all active lanes of the wave together.
__device__ __forceinline__ void filter_emit(const FilterOut& fo, bool pass, uint32_t index,
                                            float snr) {
    const uint64_t mask = __ballot(pass ? 1 : 0);
    if (mask == 0) {
        return;
    }
    const uint32_t lane   = __lane_id();
    const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(mask))) - 1U;
    uint32_t base         = 0;
    if (lane == leader) {
        base = atomicAdd(fo.count, static_cast<uint32_t>(__popcll(mask)));
    }
    base = __shfl(base, static_cast<int>(leader));
    if (pass) {
        const uint64_t below = lane == 0 ? 0ULL : (mask & ((1ULL << lane) - 1ULL));
        const uint32_t slot  = base + static_cast<uint32_t>(__popcll(below));
        if (slot < fo.capacity) {
            fo.indices[slot] = index;
            fo.snr[slot]     = snr;
        }
    }
}

/// Correctly rounded sqrtf(v) for 2^-96 <= v <= FLT_MAX without the scaling and
/// special-case handling of the general sequence: v_sqrt_f32 is within 1 ulp,
/// and the residuals v - s' * s of the two neighbours s' (exact by FMA, and
/// never strictly between two representable values of that size) pick the
/// nearest float.  Checked against the IEEE sequence for every float in range.
__device__ __forceinline__ float sqrt_rn_normal(float v) {
    const float s   = __builtin_amdgcn_sqrtf(v);
    const float sdn = __uint_as_float(__float_as_uint(s) - 1U);
    const float sup = __uint_as_float(__float_as_uint(s) + 1U);
    const float rdn = fmaf(-sdn, s, v);
    const float rup = fmaf(-sup, s, v);
    float r         = (rdn <= 0.0F) ? sdn : s;
    r               = (rup > 0.0F) ? sup : r;
    return r;
}

/// Correctly rounded 1/s for 2^-48 <= s < 2^64 (v_rcp_f32 plus one FMA
/// Newton step); checked against IEEE division for every float in range.
__device__ __forceinline__ float rcp_rn_normal(float s) {
    const float r = __builtin_amdgcn_rcpf(s);
    const float e = fmaf(-s, r, 1.0F);
    return fmaf(e, r, r);
}

/// Whether v lies in [2^-96, FLT_MAX] (excludes zero, negatives, subnormals,
/// inf and NaN), the range of sqrt_rn_normal / rcp_rn_normal.
__device__ __forceinline__ bool rn_normal_range(float v) {
    return (__float_as_uint(v) - 0x0F800000U) < (0x7F800000U - 0x0F800000U);
}

/// p = e * (1/sqrtf(v)) for four bins, bit-identical to normalise<true>: the
/// short sequences when every active lane's v is in range (wave-uniform
/// branch), the general IEEE sequence otherwise.
__device__ __forceinline__ float4 normalise4_exact(const float4 e, const float4 v) {
    const bool ok = rn_normal_range(v.x) && rn_normal_range(v.y) && rn_normal_range(v.z) &&
                    rn_normal_range(v.w);
    float4 p;
    if (__all(ok)) {
        p.x = e.x * rcp_rn_normal(sqrt_rn_normal(v.x));
        p.y = e.y * rcp_rn_normal(sqrt_rn_normal(v.y));
        p.z = e.z * rcp_rn_normal(sqrt_rn_normal(v.z));
        p.w = e.w * rcp_rn_normal(sqrt_rn_normal(v.w));
    } else {
        p.x = normalise<true>(e.x, v.x);
        p.y = normalise<true>(e.y, v.y);
        p.z = normalise<true>(e.z, v.z);
        p.w = normalise<true>(e.w, v.w);
    }
    return p;
}

}  // namespace loki_amd::kern
