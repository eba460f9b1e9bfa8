// Device code of the time brute fold shared by the fold kernels
// (k_brutefold.hip) and the batch's persistent kernel (k_tree.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "grid.hpp"
#include "kernels.hpp"
#include "storage.hpp"

namespace loki_amd::kern {
namespace {

constexpr int kFoldBlock = 256;
constexpr int kPairsBlk  = 2048;  // (freq, bin) pairs per fold block
constexpr int kLdsSamps  = 4096;  // segments up to this length are staged in LDS

// One lane per (segment, freq, bin) pair; consecutive lanes are consecutive
// bins, so list reads and output writes are contiguous.  USE_LDS: the
// segment's samples are staged in the LDS arrays s_e and s_v (kLdsSamps
// floats each, seglen <= kLdsSamps; longer segments take the run path).
// TO: stored element type of the fold (float, or _Float16 / __bf16 with each
// FP32 sum multiplied by `scale` before rounding).
template <bool USE_LDS, typename TO>
__device__ __forceinline__ void bf_fold_block(uint32_t bid, const float* __restrict__ ts_e,
                                              const float* __restrict__ ts_v, uint32_t seglen,
                                              uint32_t nfreqs, uint32_t nbins,
                                              uint32_t blocks_per_seg,
                                              const uint32_t* __restrict__ list,
                                              const uint64_t* __restrict__ base_f,
                                              const uint32_t* __restrict__ counts,
                                              TO* __restrict__ fold, float scale,
                                              float* __restrict__ s_e,
                                              float* __restrict__ s_v) {
    const uint64_t seg   = bid / blocks_per_seg;
    const uint32_t pbase = (bid % blocks_per_seg) * kPairsBlk;
    const uint32_t npair = nfreqs * nbins;
    const float* __restrict__ ge = ts_e + seg * seglen;
    const float* __restrict__ gv = ts_v + seg * seglen;
    if constexpr (USE_LDS) {
        for (uint32_t k = threadIdx.x; k < seglen; k += kFoldBlock) {
            s_e[k] = ge[k];
            s_v[k] = gv[k];
        }
        __syncthreads();
    }
    const float* __restrict__ e = USE_LDS ? s_e : ge;
    const float* __restrict__ v = USE_LDS ? s_v : gv;
    const uint32_t pend = min(pbase + static_cast<uint32_t>(kPairsBlk), npair);
    for (uint32_t p = pbase + threadIdx.x; p < pend; p += kFoldBlock) {
        const uint32_t f   = p / nbins;
        const uint32_t b   = p - f * nbins;
        const uint32_t cnt = counts[p];
        const uint32_t* __restrict__ lst = list + base_f[f] + b;
        float se = 0.0F;
        float sv = 0.0F;
        uint32_t k = 0;
        for (; k + 4 <= cnt; k += 4) {
            const uint32_t i0 = lst[static_cast<uint64_t>(k) * nbins];
            const uint32_t i1 = lst[static_cast<uint64_t>(k + 1) * nbins];
            const uint32_t i2 = lst[static_cast<uint64_t>(k + 2) * nbins];
            const uint32_t i3 = lst[static_cast<uint64_t>(k + 3) * nbins];
            const float e0 = e[i0], e1 = e[i1], e2 = e[i2], e3 = e[i3];
            const float v0 = v[i0], v1 = v[i1], v2 = v[i2], v3 = v[i3];
            se = se + e0;
            sv = sv + v0;
            se = se + e1;
            sv = sv + v1;
            se = se + e2;
            sv = sv + v2;
            se = se + e3;
            sv = sv + v3;
        }
        for (; k < cnt; ++k) {
            const uint32_t i = lst[static_cast<uint64_t>(k) * nbins];
            se = se + e[i];
            sv = sv + v[i];
        }
        TO* __restrict__ out = fold + (seg * nfreqs + f) * 2 * nbins;
        if constexpr (std::is_same_v<TO, float>) {
            out[b]         = se;
            out[nbins + b] = sv;
        } else {
            out[b]         = narrow1<TO>(se * scale);
            out[nbins + b] = narrow1<TO>(sv * scale);
        }
    }
}

// ---------------------------------------------------------------------------
// Long segments: bucket sums over contiguous runs, no lists.
// For one frequency the unwrapped bin U(k) (grid::unwrapped_bin) is
// non-decreasing in the sample index k, so bin b's samples are the runs
// [K(u), K(u + 1)) for u = b (mod nbins), K(u) = min{k : U(k) >= u}.  Each lane
// finds its run boundaries from an analytic estimate corrected against the
// exact U (a binary search if the estimate is far off) and sums the runs in
// ascending k: the same elements in the same order as the bucket list.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t u_at(uint32_t k, double tsamp, double t_ref, double f,
                                        uint32_t nbins) {
    const double t = (static_cast<double>(k) * tsamp) - t_ref;
    return grid::unwrapped_bin(t, f, nbins);
}

/// Per-frequency constants of run_bound: k_est(u) = (u - 1/2) A + B with
/// A = 1 / (nbins f tsamp) samples per bin and B = t_ref / tsamp, and the bound
/// eta (in bins) on the rounding of U(k) against nbins (k tsamp - t_ref) f:
/// the FP32 conversion and +0.5F of the bin ((2 nbins + 1/2) 2^-24) plus the
/// FP64 evaluation of the phase (3 * 2^-53 (n tsamp + |t_ref|) f nbins).
struct RunGeom {
    double f, A, B, eta_len;
};

__device__ __forceinline__ RunGeom run_geom(double f, uint32_t n, double tsamp, double t_ref,
                                            uint32_t nbins) {
    RunGeom g;
    g.f             = f;
    const double nb = static_cast<double>(nbins);
    g.A             = 1.0 / (nb * f * tsamp);
    g.B             = t_ref / tsamp;
    const double eta = ((2.0 * nb) + 0.5) * 0x1p-24 +
                       3.0 * 0x1p-53 * ((static_cast<double>(n) * tsamp) + fabs(t_ref)) * f * nb;
    g.eta_len       = eta * g.A;  // in samples
    return g;
}

/// K(u) = min{k in [0, n] : U(k) >= u} (n if none).  U(k) >= u exactly when
/// nbins (k tsamp - t_ref) f >= u - 1/2 up to eta bins, i.e. k >= k_est(u) up to
/// eta * A samples; when k_est is farther than 4x that (plus its own FP64
/// rounding) from an integer, K = ceil(k_est) needs no check.  Otherwise the
/// estimate is corrected against the exact U (a binary search if it is far).
__device__ __forceinline__ uint32_t run_bound(int64_t u, uint32_t n, double tsamp, double t_ref,
                                              const RunGeom& g, uint32_t nbins) {
    const double est = ((static_cast<double>(u) - 0.5) * g.A) + g.B;
    if (est > 0.0 && est < static_cast<double>(n)) {
        const double fr = est - floor(est);
        const double m  = (4.0 * g.eta_len) + 16.0 * 0x1p-53 * (est + 2.0 * fabs(g.B)) + 1e-9;
        if (fr > m && fr < 1.0 - m) {
            return static_cast<uint32_t>(ceil(est));
        }
    }
    const double f = g.f;
    uint32_t k = 0;
    if (est >= static_cast<double>(n)) {
        k = n;
    } else if (est > 0.0) {
        k = static_cast<uint32_t>(ceil(est));
        k = k > n ? n : k;
    }
    for (int it = 0; it < 4 && k > 0 && u_at(k - 1, tsamp, t_ref, f, nbins) >= u; ++it) {
        --k;
    }
    for (int it = 0; it < 4 && k < n && u_at(k, tsamp, t_ref, f, nbins) < u; ++it) {
        ++k;
    }
    const bool lo_ok = k == 0 || u_at(k - 1, tsamp, t_ref, f, nbins) < u;
    const bool hi_ok = k == n || u_at(k, tsamp, t_ref, f, nbins) >= u;
    if (lo_ok && hi_ok) {
        return k;
    }
    uint32_t lo = 0, hi = n;  // invariant: answer in [lo, hi]
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (u_at(mid, tsamp, t_ref, f, nbins) >= u) {
            hi = mid;
        } else {
            lo = mid + 1;
        }
    }
    return lo;
}

// Long cycles: when a cycle is longer than a chunk, a chunk holds runs of only
// a short arc of bins, so lanes bound to fixed bins mostly idle.  Here a block
// owns NF frequencies of one segment with all their bins; their running sums
// live in LDS, and per chunk the block enumerates the (frequency, bin) units
// that have samples in the chunk -- for each frequency the bins of the unwrapped
// range [U(c0), U(c1 - 1)] -- and deals them out to the lanes.  A unit adds all
// runs of its bin inside the chunk to the bin's running sum in ascending
// sample order, so every bucket is still the same ordered FP32 sum.
constexpr int kUnitChunk = 2048;  // samples staged per chunk
constexpr int kUnitMaxNF = 64;    // frequencies per block (one wave scans their unit counts)

__host__ __device__ inline uint32_t unit_nf(uint32_t nbins) {
    const uint32_t nf = 4096U / nbins;  // running sums: nf * nbins * 8 bytes <= 32 KiB
    return nf < 1U ? 1U : (nf > static_cast<uint32_t>(kUnitMaxNF) ? kUnitMaxNF : nf);
}

__host__ __device__ inline size_t unit_lds_bytes(uint32_t nbins) {
    const size_t nf = unit_nf(nbins);
    return kUnitChunk * sizeof(float2) + nf * nbins * sizeof(float2) +
           nf * (sizeof(RunGeom) + 2 * sizeof(int64_t)) + (kUnitMaxNF + 1) * sizeof(uint32_t);
}

template <typename TO>
__device__ __forceinline__ void bf_units_block(uint32_t bid, const float* __restrict__ ts_e,
                                               const float* __restrict__ ts_v, uint32_t seglen,
                                               uint32_t nfreqs, uint32_t nbins,
                                               uint32_t blocks_per_seg,
                                               const double* __restrict__ freqs, double tsamp,
                                               double t_ref, TO* __restrict__ fold, float scale) {
    extern __shared__ __align__(16) unsigned char s_mem[];
    const uint32_t NF  = unit_nf(nbins);
    auto* s_ev         = reinterpret_cast<float2*>(s_mem);  // [kUnitChunk], 16-byte aligned
    auto* acc          = s_ev + kUnitChunk;                   // [NF][nbins]
    auto* s_geom       = reinterpret_cast<RunGeom*>(acc + NF * nbins);  // [NF]
    auto* s_ulo        = reinterpret_cast<int64_t*>(s_geom + NF);        // [NF]
    auto* s_uhi        = s_ulo + NF;                                     // [NF]
    auto* s_pref       = reinterpret_cast<uint32_t*>(s_uhi + NF);        // [kUnitMaxNF + 1]
    const uint64_t seg = bid / blocks_per_seg;
    const uint32_t f0  = (bid % blocks_per_seg) * NF;
    const uint32_t nf  = min(NF, nfreqs - f0);
    const float* __restrict__ ge = ts_e + seg * seglen;
    const float* __restrict__ gv = ts_v + seg * seglen;
    const int64_t nb = nbins;
    const bool pow2  = (nbins & (nbins - 1U)) == 0;
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < nf * nbins; i += kFoldBlock) {
        acc[i] = make_float2(0.0F, 0.0F);
    }
    if (tid < nf) {
        s_geom[tid] = run_geom(freqs[f0 + tid], seglen, tsamp, t_ref, nbins);
    }
    for (uint32_t c0 = 0; c0 < seglen; c0 += kUnitChunk) {
        const uint32_t c1 = min(c0 + static_cast<uint32_t>(kUnitChunk), seglen);
        __syncthreads();  // previous chunk's units are done with s_ev and s_pref
        for (uint32_t k = c0 + tid; k < c1; k += kFoldBlock) {
            s_ev[k - c0] = make_float2(ge[k], gv[k]);
        }
        if (tid < kUnitMaxNF) {  // first wave: unit counts and their exclusive scan
            uint32_t cnt = 0;
            if (tid < nf) {
                const double f   = s_geom[tid].f;
                const int64_t lo = u_at(c0, tsamp, t_ref, f, nbins);
                const int64_t hi = u_at(c1 - 1, tsamp, t_ref, f, nbins);
                s_ulo[tid]       = lo;
                s_uhi[tid]       = hi;
                cnt = static_cast<uint32_t>(hi - lo + 1 < nb ? hi - lo + 1 : nb);
            }
            uint32_t incl = cnt;
            for (uint32_t o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o);
                incl += (tid % 64 >= o) ? y : 0U;
            }
            s_pref[tid + 1] = incl;
            if (tid == 0) {
                s_pref[0] = 0;
            }
        }
        __syncthreads();
        const uint32_t total = s_pref[nf];
        uint32_t fi = 0;  // units are ordered by frequency and t only grows
        for (uint32_t t = tid; t < total; t += kFoldBlock) {
            uint32_t lo = fi, hi = nf;  // largest fi with s_pref[fi] <= t
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (s_pref[mid] <= t) {
                    lo = mid;
                } else {
                    hi = mid;
                }
            }
            fi                 = lo;
            const int64_t u0   = s_ulo[fi] + static_cast<int64_t>(t - s_pref[fi]);
            const int64_t uhi  = s_uhi[fi];
            const RunGeom g    = s_geom[fi];
            // bin of the unit: u0 mod nbins (a mask for power-of-two nbins, which holds
            // for negative u0 too in two's complement)
            const uint32_t b   = pow2 ? static_cast<uint32_t>(u0) & (nbins - 1U)
                                      : static_cast<uint32_t>(((u0 % nb) + nb) % nb);
            float2* slot       = acc + fi * nbins + b;
            float ae = slot->x, av = slot->y;
            for (int64_t u = u0; u <= uhi; u += nb) {
                uint32_t a = run_bound(u, seglen, tsamp, t_ref, g, nbins);
                uint32_t z = run_bound(u + 1, seglen, tsamp, t_ref, g, nbins);
                a = a > c0 ? a : c0;
                z = z < c1 ? z : c1;
                uint32_t k = a;
                if ((k & 1U) != 0 && k < z) {  // c0 is even: align to a sample pair
                    const float2 x = s_ev[k - c0];
                    ae = ae + x.x;
                    av = av + x.y;
                    ++k;
                }
                // 16-byte reads of two (e, v) pairs (fewer LDS instructions, and 8 lanes
                // per LDS pass instead of 16), 16 samples per step, then 8, then one by
                // one: every sample is still added in ascending order
                const auto* __restrict__ p = reinterpret_cast<const float4*>(s_ev) +
                                              (k < z ? (k - c0) / 2 : 0);
                const uint32_t npair = k < z ? (z - k) / 2 : 0;
                const float4* const pend = p + npair;
                for (; p + 8 <= pend; p += 8) {
                    float4 y[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        y[i] = p[i];
                    }
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        ae = ae + y[i].x;
                        av = av + y[i].y;
                        ae = ae + y[i].z;
                        av = av + y[i].w;
                    }
                }
                if (p + 4 <= pend) {
                    float4 y[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        y[i] = p[i];
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        ae = ae + y[i].x;
                        av = av + y[i].y;
                        ae = ae + y[i].z;
                        av = av + y[i].w;
                    }
                    p += 4;
                }
                for (; p < pend; ++p) {
                    const float4 y = *p;
                    ae = ae + y.x;
                    av = av + y.y;
                    ae = ae + y.z;
                    av = av + y.w;
                }
                k += 2 * npair;
                if (k < z) {  // an odd last sample
                    const float2 x = s_ev[k - c0];
                    ae = ae + x.x;
                    av = av + x.y;
                }
            }
            *slot = make_float2(ae, av);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < nf * nbins; i += kFoldBlock) {
        const uint32_t fi    = i / nbins;
        const uint32_t b     = i - fi * nbins;
        TO* __restrict__ out = fold + (seg * nfreqs + f0 + fi) * 2 * nbins;
        if constexpr (std::is_same_v<TO, float>) {
            out[b]         = acc[i].x;
            out[nbins + b] = acc[i].y;
        } else {
            out[b]         = narrow1<TO>(acc[i].x * scale);
            out[nbins + b] = narrow1<TO>(acc[i].y * scale);
        }
    }
}

}  // namespace
}  // namespace loki_amd::kern
