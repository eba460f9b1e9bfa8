// Device helpers of the boxcar S/N (score.cpp:95-172, utils.cpp:22-31).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

namespace loki_amd::kern {

/// max_j (hi[j] - ps[j]) over j < NB.  fmaxf pairs lower to v_max3_f32.  For
/// non-NaN differences this equals the reference's running
/// std::max(max_diff, d) (utils.cpp:22-31); NaN differences are skipped by
/// both.  The only conceivable divergence is which zero is kept when +0.0
/// and -0.0 tie, which cannot change the S/N value.
template <int NB>
__device__ __forceinline__ float window_max(const float* __restrict__ hi, const float* ps) {
    float dmax = -FLT_MAX;
#pragma unroll
    for (int j = 0; j + 1 < NB; j += 2) {
        const float d0 = hi[j] - ps[j];
        const float d1 = hi[j + 1] - ps[j + 1];
        dmax           = __builtin_fmaxf(dmax, __builtin_fmaxf(d0, d1));
    }
    if constexpr ((NB & 1) != 0) {
        dmax = __builtin_fmaxf(dmax, hi[NB - 1] - ps[NB - 1]);
    }
    return dmax;
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

}  // namespace loki_amd::kern
