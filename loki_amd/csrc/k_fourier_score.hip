// Fused return to the time domain and boxcar S/N of the last Fourier-domain
// level: the inverse real DFT (fft_cuda.cu:357-448, 1/n normalised) of each
// final fold row runs on the matrix cores as [rows x (re, im of nbins/2+1
// harmonics)] x [(re, im) x nbins] (v_mfma_f32_16x16x4f32), the profile
// p = e / sqrt(v) is formed from the accumulators, and the S/N of
// score.cpp:95-172 is taken from a per-lane LDS prefix-sum row.  The time-domain final
// fold (2 * nbins floats per trial) is never written.
//
// One wave owns 64 consecutive trials: 8 MFMA tiles of 8 trials (16 rows:
// trial-major, e before v) whose Fourier rows are staged contiguously in LDS,
// then one lane per trial for the prefix sum and width maxima.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "kernels.hpp"
#include "score_device.hpp"

namespace loki_amd::kern {
namespace {

using f32x4 = float __attribute__((ext_vector_type(4)));

constexpr int kWaveTrials = 64;
constexpr int kTileTrials = 8;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NB> struct IrsGeom {
    static constexpr int kNh   = NB / 2 + 1;
    static constexpr int kRow  = 2 * kNh;                 // floats per component row
    static constexpr int kKs   = (kRow + 3) / 4;          // K steps of 4
    static constexpr int kNt   = NB / 16;                 // N tiles of 16 bins
    static constexpr int kTile = kTileTrials * 2 * kRow;  // floats per staged tile
};

/// Inverse real DFT basis entry B[k][n] (k = 2h + {0: re, 1: im}):
/// c_h cos(2 pi h n / NB) and -c_h sin(2 pi h n / NB), c_h = 1/NB for h = 0
/// and h = NB/2 (whose imaginary parts are ignored), 2/NB otherwise, 0 for
/// the padding k >= nbins + 2.  tw[t] = (cos, sin)(2 pi t / NB).
template <int NB>
__device__ __forceinline__ float irdft_basis(const float2* tw, int k, int n, int zero) {
    constexpr int kRow = IrsGeom<NB>::kRow;
    const int h        = k >> 1;
    const bool edge    = h == 0 || h == NB / 2;
    const bool imag    = (k & 1) != 0;
    const float c      = (k >= kRow || (imag && edge)) ? 0.0F
                         : edge                        ? 1.0F / static_cast<float>(NB)
                                                       : 2.0F / static_cast<float>(NB);
    const float2 w     = tw[((h * n) & (NB - 1)) + zero];
    return c * (imag ? -w.y : w.x);
}

template <int NB, bool FINDMAX, bool EXACT>
__global__ void __launch_bounds__(64)
k_irfft_score(const float* __restrict__ fold, uint64_t ntrials, WidthTable wt, uint32_t stride,
              float* __restrict__ scores, FilterOut fo) {
    using G = IrsGeom<NB>;
    extern __shared__ float lds[];
    float* __restrict__ stage = lds;                  // [kTile]
    float* __restrict__ prow  = lds + G::kTile;       // [64][stride]
    auto* tw = reinterpret_cast<float2*>(prow + kWaveTrials * stride);  // [NB]
    const int lane = static_cast<int>(threadIdx.x);
    const int mrow = lane & 15;   // A row / D column within the tile
    const int kq   = lane >> 4;   // A column / D row group
    for (int t = lane; t < NB; t += 64) {
        double sn, cs;
        sincospi(2.0 * static_cast<double>(t) / static_cast<double>(NB), &sn, &cs);
        tw[t] = make_float2(static_cast<float>(cs), static_cast<float>(sn));
    }
    wave_sync();
    const uint64_t nfloats = ntrials * 2 * G::kRow;
    const uint64_t ngroups = (ntrials + kWaveTrials - 1) / kWaveTrials;
    constexpr int kLoads   = (G::kTile / 4 + 63) / 64;  // float4 loads per lane per tile
    float4 pre[kLoads];
    // loads of one tile (8 trials from trial c) into registers; zeros past the end
    auto load_tile = [&](uint64_t c) {
        const uint64_t base = c * 2 * G::kRow;  // 16-byte aligned: kTile % 4 == 0
#pragma unroll
        for (int i = 0; i < kLoads; ++i) {
            const int i4     = lane + 64 * i;
            const uint64_t g = base + 4 * static_cast<uint64_t>(i4);
            float4 v         = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
            if (i4 < G::kTile / 4) {
                if (g + 4 <= nfloats) {
                    v = *reinterpret_cast<const float4*>(fold + g);
                } else {
                    v.x = g + 0 < nfloats ? fold[g + 0] : 0.0F;
                    v.y = g + 1 < nfloats ? fold[g + 1] : 0.0F;
                    v.z = g + 2 < nfloats ? fold[g + 2] : 0.0F;
                }
            }
            pre[i] = v;
        }
    };
    for (uint64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const uint64_t c0 = grp * kWaveTrials;
        load_tile(c0);
#pragma unroll 1
        for (int t = 0; t < kWaveTrials / kTileTrials; ++t) {
#pragma unroll
            for (int i = 0; i < kLoads; ++i) {
                const int i4 = lane + 64 * i;
                if (i4 < G::kTile / 4) {
                    *reinterpret_cast<float4*>(stage + 4 * i4) = pre[i];
                }
            }
            wave_sync();
            // prefetch this group's next tile while the matrix cores run
            if (t + 1 < kWaveTrials / kTileTrials) {
                load_tile(c0 + static_cast<uint64_t>(t + 1) * kTileTrials);
            }
            f32x4 acc[G::kNt];
#pragma unroll
            for (int nt = 0; nt < G::kNt; ++nt) {
                acc[nt] = f32x4{0.0F, 0.0F, 0.0F, 0.0F};
            }
            // an opaque zero keeps the basis reads inside the loop: hoisted, the 4 * kKs
            // basis values would hold that many registers for the whole kernel
            int zero = 0;
            asm volatile("" : "+v"(zero));
#pragma unroll
            for (int ks = 0; ks < G::kKs; ++ks) {
                const int k   = 4 * ks + kq;
                const float a = k < G::kRow ? stage[mrow * G::kRow + k] : 0.0F;
#pragma unroll
                for (int nt = 0; nt < G::kNt; ++nt) {
                    const float b = irdft_basis<NB>(tw, k, 16 * nt + mrow, zero);
                    acc[nt]       = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[nt], 0, 0, 0);
                }
            }
            // D rows 4*kq + v: trial 2*kq + v/2 of the tile, component v % 2
            const int tr = t * kTileTrials + 2 * kq;
#pragma unroll
            for (int nt = 0; nt < G::kNt; ++nt) {
                const int bin = 16 * nt + mrow;
                prow[tr * stride + bin]       = normalise<EXACT>(acc[nt][0], acc[nt][1]);
                prow[(tr + 1) * stride + bin] = normalise<EXACT>(acc[nt][2], acc[nt][3]);
            }
            wave_sync();  // the next tile overwrites the staged rows
        }
        const uint64_t c = c0 + static_cast<uint64_t>(lane);
        if (c < ntrials) {
            float* __restrict__ row = prow + lane * stride;
            float ps[NB];
            float acc = 0.0F;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                acc   = (j == 0) ? row[0] : acc + row[j];
                ps[j] = acc;
            }
            const float sum = ps[NB - 1];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                row[j] = ps[j];
            }
            // psum[i] for NB <= i < NB + w is psum[i - NB] + sum (the same FP32 add as a
            // stored wrap extension), read through the power-of-two index mask
            float best = -FLT_MAX;
            for (uint32_t iw = 0; iw < wt.n; ++iw) {
                const uint32_t w = wt.w[iw];
                float dmax       = -FLT_MAX;
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    const uint32_t i = static_cast<uint32_t>(j) + w;
                    float hv         = row[i & (NB - 1)];
                    hv               = i >= static_cast<uint32_t>(NB) ? hv + sum : hv;
                    dmax             = __builtin_fmaxf(dmax, hv - ps[j]);
                }
                const float snr = ((wt.h[iw] + wt.b[iw]) * dmax) - (wt.b[iw] * sum);
                if constexpr (FINDMAX) {
                    best = (best < snr) ? snr : best;
                } else if (fo.count != nullptr) {
                    filter_emit(fo, snr > fo.threshold, static_cast<uint32_t>(c * wt.n + iw),
                                snr);
                } else {
                    scores[c * wt.n + iw] = snr;
                }
            }
            if constexpr (FINDMAX) {
                if (fo.count != nullptr) {
                    filter_emit(fo, best > fo.threshold, static_cast<uint32_t>(c), best);
                } else {
                    scores[c] = best;
                }
            }
        }
        wave_sync();  // rows are rewritten by the next group
    }
}

template <int NB, bool FINDMAX, bool EXACT>
void launch_irs(const float2* fold, uint64_t ntrials, const WidthTable& wt, float* scores,
                const FilterOut& fo, hipStream_t s) {
    using G = IrsGeom<NB>;
    const uint32_t stride = NB + 1;  // odd: conflict-free per-lane rows
    const size_t lds = (static_cast<size_t>(G::kTile) + kWaveTrials * stride + 2 * NB) *
                       sizeof(float);
    LOKI_REQUIRE(lds <= 65536, "irfft+score: LDS row too long");
    const uint64_t ngroups = (ntrials + kWaveTrials - 1) / kWaveTrials;
    const uint64_t nblocks = ngroups < 8192 ? ngroups : 8192;
    k_irfft_score<NB, FINDMAX, EXACT><<<static_cast<uint32_t>(nblocks), 64, lds, s>>>(
        reinterpret_cast<const float*>(fold), ntrials, wt, stride, scores, fo);
}

template <int NB>
void dispatch_irs(const float2* fold, uint64_t ntrials, const WidthTable& wt, float* scores,
                  bool find_max, bool exact, const FilterOut& fo, hipStream_t s) {
    if (find_max) {
        return exact ? launch_irs<NB, true, true>(fold, ntrials, wt, scores, fo, s)
                     : launch_irs<NB, true, false>(fold, ntrials, wt, scores, fo, s);
    }
    return exact ? launch_irs<NB, false, true>(fold, ntrials, wt, scores, fo, s)
                 : launch_irs<NB, false, false>(fold, ntrials, wt, scores, fo, s);
}

}  // namespace

bool irfft_score_supported(uint32_t nbins) { return nbins == 16 || nbins == 32 || nbins == 64; }

void irfft_score(const float2* fold, uint64_t ntrials, uint32_t nbins, const WidthTable& wt,
                 float* scores, bool find_max, bool exact, hipStream_t s, const FilterOut* fo) {
    LOKI_REQUIRE(irfft_score_supported(nbins), "irfft+score: nbins must be 16, 32 or 64");
    if (ntrials == 0) {
        return;
    }
    const FilterOut f = fo != nullptr ? *fo : FilterOut{};
    if (fo != nullptr) {
        LOKI_REQUIRE(fo->count != nullptr, "score filter: null counter");
        LOKI_REQUIRE(ntrials * (find_max ? 1U : wt.n) <= 0xFFFFFFFFULL,
                     "score filter: trial index exceeds 32 bits");
    }
    switch (nbins) {
    case 16: return dispatch_irs<16>(fold, ntrials, wt, scores, find_max, exact, f, s);
    case 32: return dispatch_irs<32>(fold, ntrials, wt, scores, find_max, exact, f, s);
    default: return dispatch_irs<64>(fold, ntrials, wt, scores, find_max, exact, f, s);
    }
}

}  // namespace loki_amd::kern
