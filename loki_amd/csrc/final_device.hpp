// Device code of the time-domain fused last level shared by k_final.hip and
// the batch's persistent kernel (k_tree.hip): merge + normalisation of 4 bins
// and the boxcar windows of a prefix sum held in registers.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <utility>

#include "kernels.hpp"
#include "merge_device.hpp"
#include "resolve_device.hpp"
#include "score_device.hpp"

namespace loki_amd::kern {
namespace {

constexpr uint32_t kFinalWideBlock = 256;  ///< lanes of a final_tile workgroup

template <bool EXACT>
__device__ __forceinline__ float4 normalise4(const float4 e, const float4 v) {
    if constexpr (EXACT) {
        return normalise4_exact(e, v);
    } else {
        float4 p;
        p.x = normalise<false>(e.x, v.x);
        p.y = normalise<false>(e.y, v.y);
        p.z = normalise<false>(e.z, v.z);
        p.w = normalise<false>(e.w, v.w);
        return p;
    }
}

/// e = tail + head, v = tail + head (one FP32 add each), unscaled for
/// reduced-precision parents, normalised and stored as 4 profile bins at
/// row[(b0 + k) mod NB], k < 4 (b0 = 4*i4 for aligned output, or rotated by
/// the tail shift when the tail was loaded unrotated).
template <int NB, bool EXACT, typename TI>
__device__ __forceinline__ void merge_store(const float4 te, const float4 tv, const float4 he,
                                            const float4 hv, float unscale,
                                            float* __restrict__ row, uint32_t b0) {
    float4 e, v;
    e.x = te.x + he.x;
    e.y = te.y + he.y;
    e.z = te.z + he.z;
    e.w = te.w + he.w;
    v.x = tv.x + hv.x;
    v.y = tv.y + hv.y;
    v.z = tv.z + hv.z;
    v.w = tv.w + hv.w;
    if constexpr (kScaledHalf<TI>) {
        e.x *= unscale;
        e.y *= unscale;
        e.z *= unscale;
        e.w *= unscale;
        v.x *= unscale;
        v.y *= unscale;
        v.z *= unscale;
        v.w *= unscale;
    }
    const float4 p = normalise4<EXACT>(e, v);
    row[(b0 + 0) & (NB - 1)] = p.x;
    row[(b0 + 1) & (NB - 1)] = p.y;
    row[(b0 + 2) & (NB - 1)] = p.z;
    row[(b0 + 3) & (NB - 1)] = p.w;
}

/// Widest rows whose boxcar windows are formed in registers (compile-time
/// window offsets, no prefix-sum row in LDS); wider rows use the LDS row.
constexpr int kRegWindowMaxNB = 32;

/// max_j (hi_j - ps[j]) for the fixed width W, hi_j = ps[j + W] or, past the
/// end of the row, ext[j + W - NB] = ps[j + W - NB] + sum: the same operands
/// as the LDS prefix-sum row of window_max.
template <int NB, int W>
__device__ __forceinline__ float window_max_fixed(const float (&ps)[NB],
                                                  const float (&ext)[NB - 1]) {
    float ma = -FLT_MAX, mb = -FLT_MAX;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const float hi = (j + W < NB) ? ps[(j + W) % NB] : ext[(j + W - NB) % (NB - 1)];
        const float d  = hi - ps[j];
        if ((j & 2) == 0) {
            ma = __builtin_fmaxf(ma, d);
        } else {
            mb = __builtin_fmaxf(mb, d);
        }
    }
    return __builtin_fmaxf(ma, mb);
}

/// window_max_fixed<NB, w> for the runtime, wave-uniform width w in [1, NB).
template <int NB, int... Ws>
__device__ __forceinline__ float window_max_reg(std::integer_sequence<int, Ws...>,
                                                const float (&ps)[NB],
                                                const float (&ext)[NB - 1], uint32_t w) {
    float r = 0.0F;
    (void)((w == static_cast<uint32_t>(Ws + 1) ? (r = window_max_fixed<NB, Ws + 1>(ps, ext), true)
                                               : false) ||
           ...);
    return r;
}

/// Fused last level (freq-only, FP32 parents, exact normalisation, one
/// score per width) of R = 256 / WS trials with 256 lanes: final_block's
/// three steps, WS lanes scoring each trial (lane q takes widths q, q + WS,
/// ...: the boxcar-width schedule).  `tile` indexes t.ts (built with R rows
/// per tile).  LDS: R profile rows of `stride` floats, then R records.  Every
/// FP32 operation is final_block's, so the scores are bit-identical.
template <int NB, int WS>
__device__ __forceinline__ void final_tile(uint32_t tile, const FinalTask& t,
                                           unsigned char* __restrict__ lds) {
    constexpr uint32_t R     = kFinalWideBlock / WS;
    constexpr uint32_t C4    = NB / 4;
    constexpr uint32_t LPR   = C4;                // lanes per row in the merge
    constexpr uint32_t RPI   = kFinalWideBlock / LPR;  // rows per merge step
    static_assert(R % RPI == 0, "whole merge steps per tile");
    static_assert(WS == 1 || R % 64 == 0, "the lanes of one wave score the same width");
    constexpr uint32_t NSTEP = R / RPI;
    constexpr uint32_t U     = NSTEP >= 4 ? 4 : NSTEP;
    const uint32_t stride    = t.fg.stride;
    float* __restrict__ s_p  = reinterpret_cast<float*>(lds);
    auto* __restrict__ s_cd  = reinterpret_cast<FFACoord*>(lds + R * stride * sizeof(float));
    const Tile tl            = tile_of(t.ts, tile, R);
    if (tl.count == 0) {
        return;  // the whole workgroup: an idle tile of a partial run
    }
    const uint32_t tid  = threadIdx.x;
    const uint32_t nrow = tl.count;
    const float* __restrict__ in = static_cast<const float*>(t.in);

    // 1. shift tables of the tile's trials (FP64)
    for (uint32_t r = tid; r < nrow; r += kFinalWideBlock) {
        s_cd[r] = resolve_pair<1>(t.lv, t.fg.dv, tl.c0 + r);
    }
    __syncthreads();

    // 2. merge + normalise: lane (rsub, i4) forms bins 4*i4 .. 4*i4+3 of rows
    //    rsub, rsub + RPI, ... (the tail is unshifted for freq-only levels)
    {
        const uint32_t rsub = tid / LPR;
        const uint32_t i4   = tid % LPR;
#pragma unroll
        for (uint32_t s0 = 0; s0 < NSTEP; s0 += U) {
            float4 te[U], tv[U], he[U], hv[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t r = rsub + (s0 + u) * RPI;
                if (r < nrow) {
                    const FFACoord cd = s_cd[r];
                    const auto* __restrict__ trow = reinterpret_cast<const float4*>(
                        in + static_cast<uint64_t>(cd.i_tail) * (2 * NB));
                    const auto* __restrict__ hrow = reinterpret_cast<const float4*>(
                        in + (t.nprev + static_cast<uint64_t>(cd.i_head)) * (2 * NB));
                    const uint32_t Sh = shift_int(cd.shift_head, NB);
                    te[u] = chunk4_st<float>(trow, i4, 1.0F);
                    tv[u] = chunk4_st<float>(trow + C4, i4, 1.0F);
                    he[u] = rotated4_st<NB, float>(hrow, i4, Sh, 1.0F);
                    hv[u] = rotated4_st<NB, float>(hrow + C4, i4, Sh, 1.0F);
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t r = rsub + (s0 + u) * RPI;
                if (r < nrow) {
                    merge_store<NB, true, float>(te[u], tv[u], he[u], hv[u], t.unscale,
                                                 s_p + r * stride, 4 * i4);
                }
            }
        }
    }
    __syncthreads();

    // 3. S/N: lanes q * R + row score trial `row`, widths q, q + WS, ...
    //    (lane groups of R >= 64 keep each wave on one width)
    const uint32_t row  = tid % R;
    const uint32_t q    = tid / R;
    const bool active   = row < nrow;
    float* __restrict__ rowp = s_p + row * stride;
    const uint64_t c    = tl.c0 + row;
    if constexpr (NB <= kRegWindowMaxNB) {
        // circular prefix sum and its wrap in registers, windows with
        // compile-time offsets (final_block's register path)
        if (!active) {
            return;
        }
        float ps[NB];
        float acc = 0.0F;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const float x = rowp[j];
            acc           = (j == 0) ? x : acc + x;
            ps[j]         = acc;
        }
        const float sum = ps[NB - 1];
        float ext[NB - 1];
#pragma unroll
        for (int k = 0; k < NB - 1; ++k) {
            ext[k] = ps[k] + sum;
        }
        for (uint32_t iw = q; iw < t.wt.n; iw += WS) {
            __asm__ volatile("" : "+v"(ps[0]));  // keep the windows inside the width loop
            const float dmax = window_max_reg<NB>(std::make_integer_sequence<int, NB - 1>{}, ps,
                                                  ext, t.wt.w[iw]);
            t.scores[c * t.wt.n + iw] =
                ((t.wt.h[iw] + t.wt.b[iw]) * dmax) - (t.wt.b[iw] * sum);
        }
    } else {
        // wider rows: the row's first lane group turns the profile row into
        // its prefix sum in place (the same sequential FP32 adds) followed by
        // the wrap ps[j] + sum (n + wmax slots); every window then reads both
        // of its operands from LDS, so no lane holds the row in registers
        if (active && q == 0) {
            float acc = 0.0F;
#pragma unroll 8
            for (int j = 0; j < NB; ++j) {
                const float x = rowp[j];
                acc           = (j == 0) ? x : acc + x;
                rowp[j]       = acc;
            }
            for (uint32_t j = 0; j < t.wt.wmax; ++j) {
                rowp[NB + j] = rowp[j] + acc;
            }
        }
        __syncthreads();
        if (!active) {
            return;
        }
        const float sum = rowp[NB - 1];
        for (uint32_t iw = q; iw < t.wt.n; iw += WS) {
            const float* __restrict__ hi = rowp + t.wt.w[iw];
            // window_max's two running maxima, both operands from the row
            float ma = -FLT_MAX, mb = -FLT_MAX;
#pragma unroll 4
            for (int j = 0; j < NB; j += 4) {
                ma = __builtin_fmaxf(__builtin_fmaxf(ma, hi[j] - rowp[j]), hi[j + 1] - rowp[j + 1]);
                mb = __builtin_fmaxf(__builtin_fmaxf(mb, hi[j + 2] - rowp[j + 2]),
                                     hi[j + 3] - rowp[j + 3]);
            }
            const float dmax = __builtin_fmaxf(ma, mb);
            t.scores[c * t.wt.n + iw] =
                ((t.wt.h[iw] + t.wt.b[iw]) * dmax) - (t.wt.b[iw] * sum);
        }
    }
}

}  // namespace
}  // namespace loki_amd::kern
