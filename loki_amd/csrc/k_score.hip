// Boxcar matched-filter S/N (score.cpp:95-172, utils.cpp:22-67).
//
// One lane per profile.  The lane normalises its profile (p = e * (1/sqrt(v)),
// the CPU's order), forms the circular prefix sum sequentially -- the exact
// FP32 order of the reference CPU -- and keeps it in VGPRs (statically
// indexed) plus a per-lane LDS row for the dynamically offset psum[j + w]
// reads (odd row stride: conflict-free).  Then for every width
//   snr_w = (h_w + b_w) * max_j(psum[j + w] - psum[j]) - b_w * sum.
// With -ffp-contract=off every value is bit-identical to the CPU oracle.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "kernels.hpp"
#include "score_device.hpp"

namespace loki_amd::kern {
namespace {

constexpr int kLanes = 64;

__host__ __device__ inline uint32_t psum_stride(uint32_t nbins, uint32_t wmax) {
    return (nbins + wmax) | 1U;
}

template <int NB, bool IS3D, bool FINDMAX>
__global__ void __launch_bounds__(kLanes)
k_snr_reg(const float* __restrict__ folds, uint64_t nprof, WidthTable wt, float inv_stdnoise,
          float* __restrict__ scores, FilterOut fo) {
    extern __shared__ float s_ps[];
    const uint64_t p = static_cast<uint64_t>(blockIdx.x) * kLanes + threadIdx.x;
    if (p >= nprof) {
        return;  // lanes never synchronise: each owns its LDS row
    }
    const uint32_t stride = psum_stride(NB, wt.wmax);
    float* __restrict__ lds = s_ps + threadIdx.x * stride;
    float ps[NB];
    float acc = 0.0F;
    if constexpr (IS3D) {
        const float4* __restrict__ row = reinterpret_cast<const float4*>(folds + p * 2 * NB);
#pragma unroll
        for (int i4 = 0; i4 < NB / 4; ++i4) {
            const float4 e = row[i4];
            const float4 v = row[NB / 4 + i4];
            const float x0 = e.x * (1.0F / sqrtf(v.x));
            const float x1 = e.y * (1.0F / sqrtf(v.y));
            const float x2 = e.z * (1.0F / sqrtf(v.z));
            const float x3 = e.w * (1.0F / sqrtf(v.w));
            acc = (i4 == 0) ? x0 : acc + x0;
            ps[4 * i4 + 0] = acc;
            acc = acc + x1;
            ps[4 * i4 + 1] = acc;
            acc = acc + x2;
            ps[4 * i4 + 2] = acc;
            acc = acc + x3;
            ps[4 * i4 + 3] = acc;
        }
    } else {
        const float4* __restrict__ row = reinterpret_cast<const float4*>(folds + p * NB);
#pragma unroll
        for (int i4 = 0; i4 < NB / 4; ++i4) {
            const float4 x = row[i4];
            acc = (i4 == 0) ? x.x : acc + x.x;
            ps[4 * i4 + 0] = acc;
            acc = acc + x.y;
            ps[4 * i4 + 1] = acc;
            acc = acc + x.z;
            ps[4 * i4 + 2] = acc;
            acc = acc + x.w;
            ps[4 * i4 + 3] = acc;
        }
    }
    const float sum = ps[NB - 1];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        lds[j] = ps[j];
    }
    // Wrap: psum[i] = psum[i - n] + sum for n <= i < n + wmax (wmax < n).
    for (uint32_t j = 0; j < wt.wmax; ++j) {
        lds[NB + j] = lds[j] + sum;
    }
    float best = -FLT_MAX;
    for (uint32_t iw = 0; iw < wt.n; ++iw) {
        const float dmax = window_max<NB>(lds + wt.w[iw], ps);
        const float snr = (((wt.h[iw] + wt.b[iw]) * dmax) - (wt.b[iw] * sum)) * inv_stdnoise;
        if constexpr (FINDMAX) {
            best = (best < snr) ? snr : best;
        } else if (fo.count != nullptr) {
            filter_emit(fo, snr > fo.threshold, static_cast<uint32_t>(p * wt.n + iw), snr);
        } else {
            scores[p * wt.n + iw] = snr;
        }
    }
    if constexpr (FINDMAX) {
        if (fo.count != nullptr) {
            filter_emit(fo, best > fo.threshold, static_cast<uint32_t>(p), best);
        } else {
            scores[p] = best;
        }
    }
}

// Any nbins: psum lives entirely in the lane's LDS row.
template <bool IS3D, bool FINDMAX>
__global__ void __launch_bounds__(kLanes)
k_snr_lds(const float* __restrict__ folds, uint64_t nprof, uint32_t nbins, WidthTable wt,
          float inv_stdnoise, float* __restrict__ scores, FilterOut fo) {
    extern __shared__ float s_ps[];
    const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= nprof) {
        return;
    }
    const uint32_t stride = psum_stride(nbins, wt.wmax);
    float* __restrict__ lds = s_ps + threadIdx.x * stride;
    float acc = 0.0F;
    for (uint32_t j = 0; j < nbins; ++j) {
        float x;
        if constexpr (IS3D) {
            const float* row = folds + p * 2 * nbins;
            x = row[j] * (1.0F / sqrtf(row[nbins + j]));
        } else {
            x = folds[p * nbins + j];
        }
        acc    = (j == 0) ? x : acc + x;
        lds[j] = acc;
    }
    const float sum = acc;
    for (uint32_t j = 0; j < wt.wmax; ++j) {
        lds[nbins + j] = lds[j] + sum;
    }
    float best = -FLT_MAX;
    for (uint32_t iw = 0; iw < wt.n; ++iw) {
        const uint32_t w = wt.w[iw];
        float dmax       = -FLT_MAX;
        for (uint32_t j = 0; j < nbins; ++j) {
            dmax = __builtin_fmaxf(dmax, lds[j + w] - lds[j]);  // see window_max
        }
        const float snr = (((wt.h[iw] + wt.b[iw]) * dmax) - (wt.b[iw] * sum)) * inv_stdnoise;
        if constexpr (FINDMAX) {
            best = (best < snr) ? snr : best;
        } else if (fo.count != nullptr) {
            filter_emit(fo, snr > fo.threshold, static_cast<uint32_t>(p * wt.n + iw), snr);
        } else {
            scores[p * wt.n + iw] = snr;
        }
    }
    if constexpr (FINDMAX) {
        if (fo.count != nullptr) {
            filter_emit(fo, best > fo.threshold, static_cast<uint32_t>(p), best);
        } else {
            scores[p] = best;
        }
    }
}

template <int NB, bool IS3D, bool FINDMAX>
void launch_reg(const float* folds, uint64_t nprof, const WidthTable& wt, float inv,
                float* scores, const FilterOut& fo, hipStream_t s) {
    const uint64_t blocks = (nprof + kLanes - 1) / kLanes;
    LOKI_REQUIRE(blocks < (1ULL << 31), "score grid too large");
    const size_t lds = static_cast<size_t>(kLanes) * psum_stride(NB, wt.wmax) * sizeof(float);
    k_snr_reg<NB, IS3D, FINDMAX><<<static_cast<unsigned>(blocks), kLanes, lds, s>>>(
        folds, nprof, wt, inv, scores, fo);
}

template <bool IS3D, bool FINDMAX>
void launch(const float* folds, uint64_t nprof, uint32_t nbins, const WidthTable& wt, float inv,
            float* scores, const FilterOut& fo, hipStream_t s) {
    const bool reg_fits =
        static_cast<size_t>(kLanes) * psum_stride(nbins, wt.wmax) * sizeof(float) <= 65536;
    switch (reg_fits ? nbins : 0U) {
    case 8: return launch_reg<8, IS3D, FINDMAX>(folds, nprof, wt, inv, scores, fo, s);
    case 16: return launch_reg<16, IS3D, FINDMAX>(folds, nprof, wt, inv, scores, fo, s);
    case 32: return launch_reg<32, IS3D, FINDMAX>(folds, nprof, wt, inv, scores, fo, s);
    case 64: return launch_reg<64, IS3D, FINDMAX>(folds, nprof, wt, inv, scores, fo, s);
    case 128: return launch_reg<128, IS3D, FINDMAX>(folds, nprof, wt, inv, scores, fo, s);
    default: break;
    }
    const uint32_t stride = psum_stride(nbins, wt.wmax);
    uint32_t lanes        = kLanes;
    while (lanes > 1 && static_cast<size_t>(lanes) * stride * sizeof(float) > 65536) {
        lanes /= 2;
    }
    const size_t lds      = static_cast<size_t>(lanes) * stride * sizeof(float);
    LOKI_REQUIRE(lds <= 65536, "score: nbins too large for the LDS prefix sum");
    const uint64_t blocks = (nprof + lanes - 1) / lanes;
    LOKI_REQUIRE(blocks < (1ULL << 31), "score grid too large");
    k_snr_lds<IS3D, FINDMAX><<<static_cast<unsigned>(blocks), lanes, lds, s>>>(
        folds, nprof, nbins, wt, inv, scores, fo);
}

}  // namespace

WidthTable make_width_table(const SizeType* widths, SizeType nw, SizeType nbins) {
    LOKI_REQUIRE(nw >= 1, "widths cannot be empty");
    LOKI_REQUIRE(nw <= WidthTable::kMax, "at most 64 boxcar widths are supported");
    WidthTable wt{};
    wt.n    = static_cast<uint32_t>(nw);
    wt.wmax = 0;
    for (SizeType i = 0; i < nw; ++i) {
        const SizeType w = widths[i];
        // The reference yields h = 0, b = 0/0 (NaN) for w >= nbins
        // (score.cpp:112-116); such widths are rejected here.
        LOKI_REQUIRE(w >= 1 && w < nbins, "boxcar widths must satisfy 1 <= w < nbins");
        wt.w[i] = static_cast<uint32_t>(w);
        wt.h[i] = std::sqrt(static_cast<float>(nbins - w) / static_cast<float>(nbins * w));
        wt.b[i] = static_cast<float>(w) * wt.h[i] / static_cast<float>(nbins - w);
        wt.hb[i] = wt.h[i] + wt.b[i];
        wt.wmax = std::max(wt.wmax, wt.w[i]);
    }
    return wt;
}

void snr_boxcar_3d(const float* folds, uint64_t nprof, uint32_t nbins, const WidthTable& wt,
                   float* scores, bool find_max, hipStream_t s, const FilterOut* fo) {
    if (nprof == 0) {
        return;
    }
    const FilterOut f = fo != nullptr ? *fo : FilterOut{};
    if (fo != nullptr) {
        LOKI_REQUIRE(fo->count != nullptr && (fo->capacity == 0 ||
                                              (fo->indices != nullptr && fo->snr != nullptr)),
                     "score filter: null output");
        LOKI_REQUIRE(nprof * (find_max ? 1U : wt.n) <= 0xFFFFFFFFULL,
                     "score filter: trial index exceeds 32 bits");
    }
    if (find_max) {
        launch<true, true>(folds, nprof, nbins, wt, 1.0F, scores, f, s);
    } else {
        launch<true, false>(folds, nprof, nbins, wt, 1.0F, scores, f, s);
    }
}

void snr_boxcar_2d_max(const float* folds, uint64_t nprof, uint32_t nbins, const WidthTable& wt,
                       float stdnoise, float* scores, hipStream_t s) {
    if (nprof == 0) {
        return;
    }
    launch<false, true>(folds, nprof, nbins, wt, 1.0F / stdnoise, scores, FilterOut{}, s);
}

namespace {
__global__ void k_zero_u32(uint32_t* p) { p[threadIdx.x] = 0U; }
}  // namespace

void zero_counter(uint32_t* count, hipStream_t s) {
    LOKI_REQUIRE(count != nullptr, "counter is null");
    k_zero_u32<<<1, 1, 0, s>>>(count);
}

namespace {
constexpr float kRcpLo = 0x1p-48F;  // rcp_rn_normal's range [2^-48, 2^64)
constexpr float kRcpHi = 0x1p64F;

__global__ void __launch_bounds__(256)
    k_selftest_exact_norm_sweep(uint32_t lo, uint32_t hi, unsigned long long* counts,
                                uint32_t* first_bad) {
    unsigned long long bad_s = 0, bad_r = 0, n_s = 0, n_r = 0;
    uint32_t first_s = 0xFFFFFFFFU, first_r = 0xFFFFFFFFU;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t u = lo + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; u < hi;
         u += stride) {
        const uint32_t b = static_cast<uint32_t>(u);
        const float x    = __uint_as_float(b);
        if (rn_normal_range(x)) {
            ++n_s;
            if (__float_as_uint(sqrt_rn_normal(x)) != __float_as_uint(sqrtf(x))) {
                ++bad_s;
                first_s = min(first_s, b);
            }
        }
        if (x >= kRcpLo && x < kRcpHi) {
            ++n_r;
            if (__float_as_uint(rcp_rn_normal(x)) != __float_as_uint(1.0F / x)) {
                ++bad_r;
                first_r = min(first_r, b);
            }
        }
    }
    atomicAdd(&counts[2], n_s);
    atomicAdd(&counts[3], n_r);
    if (bad_s != 0) {
        atomicAdd(&counts[0], bad_s);
        atomicMin(&first_bad[0], first_s);
    }
    if (bad_r != 0) {
        atomicAdd(&counts[1], bad_r);
        atomicMin(&first_bad[1], first_r);
    }
}

__global__ void __launch_bounds__(256)
    k_selftest_exact_norm_values(const uint32_t* bits, uint64_t n, float* out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const float x  = __uint_as_float(bits[i]);
    out[4 * i]     = sqrtf(x);
    out[4 * i + 1] = 1.0F / x;
    out[4 * i + 2] = sqrt_rn_normal(x);
    out[4 * i + 3] = rcp_rn_normal(x);
}
}  // namespace

void selftest_exact_norm_sweep(uint32_t bits_lo, uint32_t bits_hi, unsigned long long* counts,
                               uint32_t* first_bad, hipStream_t s) {
    LOKI_REQUIRE(counts != nullptr && first_bad != nullptr, "null argument");
    if (bits_hi <= bits_lo) {
        return;
    }
    k_selftest_exact_norm_sweep<<<8192, 256, 0, s>>>(bits_lo, bits_hi, counts, first_bad);
}

void selftest_exact_norm_values(const uint32_t* bits, uint64_t n, float* out, hipStream_t s) {
    LOKI_REQUIRE(bits != nullptr && out != nullptr, "null argument");
    if (n == 0) {
        return;
    }
    const uint64_t blocks = (n + 255) / 256;
    LOKI_REQUIRE(blocks <= 0x7FFFFFFFULL, "too many values");
    k_selftest_exact_norm_values<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(bits, n, out);
}

}  // namespace loki_amd::kern
