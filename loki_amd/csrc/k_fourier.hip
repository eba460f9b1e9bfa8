// Fourier-domain FFA kernels: the DFT brute fold, the phase-rotate merge and
// the small real transforms between the time and Fourier folds.
//
// Brute fold (kernels.cpp:582-720, 1156-1242; kernels_cuda.cu:780-965):
//   fold[s, f, q, m] = sum_k x_q[s*seglen + k] * exp(-2 pi i m f (k*tsamp - t_ref))
// for harmonics m = 0..nbins/2.  For one trial frequency f this is a real
// matrix product over the samples of each segment,
//   C_f[(s, q), n] = sum_k X[(s, q), k] * B_f[k, n],
//   B_f[k, 2(m-1)] = cos(2 pi m phi_k),  B_f[k, 2(m-1)+1] = -sin(2 pi m phi_k),
// phi_k = frac(f (k*tsamp - t_ref)), so it runs on the matrix cores
// (v_mfma_f32_16x16x4f32, FP32 in and out).  A workgroup owns one frequency
// and 128..1024 rows (segment, component); the twiddle block B_f is generated
// per chunk of 16 samples in LDS from FP64 phases and shared by all rows of
// the workgroup, so the transcendental work is ~1/256 of the MACs.  The
// harmonic m = 0 (the segment sum, identical for every frequency) is a
// separate FP64 reduction.
//
// Merge (kernels.cpp:241-453, 942-1128; kernels_cuda.cu:512-659):
//   out[s, c, q, k] = in[2s, i_t, q, k] w(s_t)^k + in[2s+1, i_h, q, k] w(s_h)^k,
//   w(x)^k = exp(-2 pi i k x / nbins) with the fractional shift x,
// (freq-only: tail factor 1).  One lane per (segment, coordinate, harmonic)
// handles both components; the angle k*x/nbins is reduced mod 1 in FP64 and
// the twiddle comes from an accurate sincospi plus a first-order correction
// for the FP32 rounding of the reduced angle.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"

namespace loki_amd::kern {
namespace {

using f32x4 = float __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kKc    = 16;  // samples per LDS twiddle chunk (4 MFMA k-steps)
constexpr uint32_t kBlockChunks = 8;  // chunks per matrix-core partial sum (128 samples)
constexpr int kDftBlock = 512;        // DFT workgroup: 8 waves
constexpr int kDftWaves = kDftBlock / 64;

/// 16-row blocks per wave of the DFT kernel: 8 accumulator tiles per wave
/// (at least one row block).
constexpr int dft_mrb(int ncb) { return ncb >= 8 ? 1 : 8 / ncb; }

/// exp(-2 pi i d) for d = frac of an FP64 angle in turns, accurate to about
/// one FP32 ulp: sincospi of the FP32 head of d folded to [-1/2, 1/2], then a
/// first-order correction with the FP32 tail of d.
__device__ __forceinline__ float2 twiddle_turns(double d) {
    d -= floor(d);
    if (d > 0.5) {
        d -= 1.0;
    }
    const float hi = static_cast<float>(d);
    const float lo = static_cast<float>(d - static_cast<double>(hi));
    float sn, cs;
    sincospif(2.0F * hi, &sn, &cs);
    const float dl = 6.28318530717958647692F * lo;
    const float c  = fmaf(-dl, sn, cs);
    const float s  = fmaf(dl, cs, sn);
    return make_float2(c, -s);
}

// ---------------------------------------------------------------------------
// Brute fold: harmonic 0 (segment sums, FP64 accumulation)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_bf_fourier_dc(const float* __restrict__ ts_e, const float* __restrict__ ts_v,
                float2* __restrict__ fold, uint32_t seglen, uint32_t nfreqs, uint32_t nbf) {
    __shared__ double red[2][kWaves];
    const uint64_t s = blockIdx.x;
    const float* e   = ts_e + s * seglen;
    const float* v   = ts_v + s * seglen;
    double se = 0.0, sv = 0.0;
    for (uint32_t k = threadIdx.x; k < seglen; k += kBlock) {
        se += static_cast<double>(e[k]);
        sv += static_cast<double>(v[k]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        se += __shfl_xor(se, o);
        sv += __shfl_xor(sv, o);
    }
    const uint32_t w = threadIdx.x / 64;
    if ((threadIdx.x & 63U) == 0) {
        red[0][w] = se;
        red[1][w] = sv;
    }
    __syncthreads();
    double te = 0.0, tv = 0.0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
        te += red[0][i];
        tv += red[1][i];
    }
    const float2 de = make_float2(static_cast<float>(te), 0.0F);
    const float2 dv = make_float2(static_cast<float>(tv), 0.0F);
    for (uint32_t f = threadIdx.x; f < nfreqs; f += kBlock) {
        float2* row = fold + ((s * nfreqs + f) * 2) * nbf;
        row[0]      = de;
        row[nbf]    = dv;
    }
}

// ---------------------------------------------------------------------------
// Brute fold: harmonics 1..nbf-1 on the matrix cores
// ---------------------------------------------------------------------------
// NCB: 16-column blocks (columns = (re, im) of harmonics 1..nh, nh = nbf-1);
// MRB: 16-row blocks per wave (MRB * NCB = 8 accumulator tiles per wave, so
// that the partial and running sums fit in registers at two waves per SIMD).
// Row r = 2 s + q of the workgroup's range; X[r, k] = ts_q[s*seglen + k].
//
// Lane l of an MFMA 16x16x4 holds A[l%16][l/16], B[l/16][l%16] and
// D[4*(l/16) + v][l%16] (v = 0..3).  The k order inside a chunk is permuted
// so a lane's four A values come from one float4 load: MFMA step i uses
// samples {k0 + 4g + i : g = 0..3}; B uses the same permutation.
template <int NCB>
__global__ void __launch_bounds__(kDftBlock)
k_bf_fourier_mfma(const float* __restrict__ ts_e, const float* __restrict__ ts_v,
                  float2* __restrict__ fold, const double* __restrict__ freqs, uint32_t nfreqs,
                  uint32_t nseg, uint32_t seglen, uint32_t nbf, double tsamp, double t_ref) {
    constexpr int MRB    = dft_mrb(NCB);
    constexpr int NCOLS  = 16 * NCB;
    constexpr int BSTR   = NCOLS + 4;  // LDS row stride (bank spread of the 4 k-groups)
    constexpr int WROWS  = 16 * MRB;   // rows per wave
    constexpr int GROWS  = WROWS * kDftWaves;
    __shared__ float bs[2][kKc * BSTR];

    const uint32_t nrows = 2 * nseg;
    const uint32_t fi    = blockIdx.x % nfreqs;
    const uint32_t rblk  = blockIdx.x / nfreqs;
    const uint32_t lane  = threadIdx.x & 63U;
    const uint32_t wave  = threadIdx.x >> 6;
    const uint32_t g     = lane >> 4;    // k-group of the lane
    const uint32_t li    = lane & 15U;   // row (A) / column (B, D) within a tile
    const uint32_t nh    = nbf - 1;
    const uint32_t ncols = 2 * nh;
    const double f       = freqs[fi];
    const bool vec_ok    = (seglen % kKc) == 0 &&
                        ((reinterpret_cast<uintptr_t>(ts_e) | reinterpret_cast<uintptr_t>(ts_v)) &
                         15U) == 0;

    for (int i = threadIdx.x; i < 2 * kKc * BSTR; i += kDftBlock) {
        (&bs[0][0])[i] = 0.0F;
    }

    // Rows of this lane (A operand): row = base + rb*16 + li
    const uint32_t row_base = rblk * GROWS + wave * WROWS;
    const float* xrow[MRB];
#pragma unroll
    for (int rb = 0; rb < MRB; ++rb) {
        const uint32_t r = row_base + rb * 16 + li;
        const uint32_t rr = r < nrows ? r : 0;
        xrow[rb] = ((rr & 1U) ? ts_v : ts_e) + static_cast<uint64_t>(rr >> 1) * seglen;
    }

    auto load_x = [&](uint32_t k0, f32x4* xa) {
#pragma unroll
        for (int rb = 0; rb < MRB; ++rb) {
            const uint32_t r = row_base + rb * 16 + li;
            const uint32_t k = k0 + 4 * g;
            f32x4 v = {0.0F, 0.0F, 0.0F, 0.0F};
            if (r < nrows) {
                if (vec_ok) {
                    v = *reinterpret_cast<const f32x4*>(xrow[rb] + k);
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        v[i] = (k + i < seglen) ? xrow[rb][k + i] : 0.0F;
                    }
                }
            }
            xa[rb] = v;
        }
    };

    // Twiddle generation: (kk, m) pairs of the chunk, 2 floats each.
    auto gen_b = [&](uint32_t k0, float* b) {
        const uint32_t npairs = kKc * nh;
        for (uint32_t p = threadIdx.x; p < npairs; p += kDftBlock) {
            const uint32_t kk = p / nh;
            const uint32_t m  = p % nh + 1;
            const double ph   = f * ((static_cast<double>(k0 + kk) * tsamp) - t_ref);
            const double fr   = ph - floor(ph);
            const float2 w    = twiddle_turns(fr * static_cast<double>(m));
            *reinterpret_cast<float2*>(b + kk * BSTR + 2 * (m - 1)) = w;
        }
    };

    // Two-level sum: the matrix cores accumulate blocks of kBlockChunks
    // chunks (128 samples), which are then added into `tot`.  Coherent sums
    // (an all-positive ts_v at low f*m) grow linearly, and a single FP32 sum
    // over K terms would carry ~u*sqrt(K)*sum|x|.
    f32x4 acc[MRB][NCB], tot[MRB][NCB];
#pragma unroll
    for (int rb = 0; rb < MRB; ++rb) {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            acc[rb][cb] = f32x4{0.0F, 0.0F, 0.0F, 0.0F};
            tot[rb][cb] = f32x4{0.0F, 0.0F, 0.0F, 0.0F};
        }
    }

    const uint32_t nchunks = (seglen + kKc - 1) / kKc;
    __syncthreads();  // zero-fill of bs done
    gen_b(0, bs[0]);
    f32x4 xa[MRB];
    load_x(0, xa);
    __syncthreads();
    for (uint32_t ch = 0; ch < nchunks; ++ch) {
        const float* b = bs[ch & 1U];
        f32x4 xn[MRB];
        const bool more = ch + 1 < nchunks;
        if (more) {
            load_x((ch + 1) * kKc, xn);
            gen_b((ch + 1) * kKc, bs[(ch + 1) & 1U]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float bv[NCB];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                bv[cb] = b[(4 * g + i) * BSTR + cb * 16 + li];
            }
#pragma unroll
            for (int rb = 0; rb < MRB; ++rb) {
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) {
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[rb][i], bv[cb],
                                                                      acc[rb][cb], 0, 0, 0);
                }
            }
        }
        if (more) {
#pragma unroll
            for (int rb = 0; rb < MRB; ++rb) {
                xa[rb] = xn[rb];
            }
        }
        if ((ch + 1) % kBlockChunks == 0 || !more) {
#pragma unroll
            for (int rb = 0; rb < MRB; ++rb) {
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) {
                    tot[rb][cb] += acc[rb][cb];
                    acc[rb][cb] = f32x4{0.0F, 0.0F, 0.0F, 0.0F};
                }
            }
        }
        __syncthreads();
    }

    // D[4g + v][li] of tile (rb, cb): row r = row_base + rb*16 + 4g + v,
    // column n = cb*16 + li -> harmonic n/2 + 1, (re, im)[n % 2].
#pragma unroll
    for (int rb = 0; rb < MRB; ++rb) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint32_t r = row_base + rb * 16 + 4 * g + v;
            if (r >= nrows) {
                continue;
            }
            const uint64_t s = r >> 1;
            const uint32_t q = r & 1U;
            float* out = reinterpret_cast<float*>(fold + ((s * nfreqs + fi) * 2 + q) * nbf) + 2;
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const uint32_t n = cb * 16 + li;
                if (n < ncols) {
                    out[n] = tot[rb][cb][v];
                }
            }
        }
    }
}

// Split-bf16 variant: X = Xh + Xl and B = Bh + Bl with bf16 parts (the low
// part is the bf16 rounding of the FP32 remainder, so each sum carries ~16
// significant bits), and X B ~ Xh Bh + Xh Bl + Xl Bh on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation: three bf16 products per
// term at 16x the FP32 matrix rate.  Per product the dropped and unrepresented
// terms are ~2^-17 relative.  Lane l of a 16x16x32 MFMA holds
// A[l%16][8(l/16) + j] and B[8(l/16) + j][l%16] (j = 0..7) and the same D
// layout as the FP32 kernel.  Chunks of 64 samples (two k-steps); the twiddle
// block is kept in LDS as bf16 parts, column-major with 8 consecutive samples
// per 16 bytes.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int kSteps3 = 2;             // MFMA k-steps (32 samples each) per chunk
constexpr int kKc3    = 32 * kSteps3;  // samples per chunk
constexpr int kKs3    = kKc3 + 8;      // bf16 per LDS column (16-byte aligned, bank spread)
constexpr uint32_t kBlockChunks3 = 128 / kKc3;  // chunks per matrix-core partial sum
constexpr int kDft3Block = 256;  // 4 waves: two workgroups per CU overlap their barriers
constexpr int kDft3Waves = kDft3Block / 64;

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

/// Two values at once (packed conversions, v_cvt_pk_bf16_f32): high parts and
/// the bf16 roundings of the exact FP32 remainders.
__device__ __forceinline__ void split_bf16x2(f32x2 x, bf16x2& hi, bf16x2& lo) {
    hi = __builtin_convertvector(x, bf16x2);
    const f32x2 back = __builtin_convertvector(hi, f32x2);
    lo = __builtin_convertvector(x - back, bf16x2);
}

template <int NCB>
__global__ void __launch_bounds__(kDft3Block)
k_bf_fourier_mfma3(const float* __restrict__ ts_e, const float* __restrict__ ts_v,
                   float2* __restrict__ fold, const double* __restrict__ freqs, uint32_t nfreqs,
                   uint32_t nseg, uint32_t seglen, uint32_t nbf, double tsamp, double t_ref) {
    constexpr int MRB   = dft_mrb(NCB);
    constexpr int NCOLS = 16 * NCB;
    constexpr int WROWS = 16 * MRB;
    constexpr int GROWS = WROWS * kDft3Waves;
    constexpr int TBL   = NCOLS * kKs3;  // bf16 per table
    __shared__ __align__(16) __bf16 bt[2][2][TBL];  // [buffer][hi, lo][col][k]

    const uint32_t nrows = 2 * nseg;
    const uint32_t fi    = blockIdx.x % nfreqs;
    const uint32_t rblk  = blockIdx.x / nfreqs;
    const uint32_t lane  = threadIdx.x & 63U;
    const uint32_t wave  = threadIdx.x >> 6;
    const uint32_t g     = lane >> 4;
    const uint32_t li    = lane & 15U;
    const uint32_t nh    = nbf - 1;
    const uint32_t ncols = 2 * nh;
    const double f       = freqs[fi];
    const bool vec_ok    = (seglen % kKc3) == 0 &&
                        ((reinterpret_cast<uintptr_t>(ts_e) | reinterpret_cast<uintptr_t>(ts_v)) &
                         15U) == 0;

    for (int i = threadIdx.x; i < 2 * 2 * TBL; i += kDft3Block) {
        (&bt[0][0][0])[i] = static_cast<__bf16>(0.0F);
    }

    const uint32_t row_base = rblk * GROWS + wave * WROWS;
    const float* xrow[MRB];
#pragma unroll
    for (int rb = 0; rb < MRB; ++rb) {
        const uint32_t r  = row_base + rb * 16 + li;
        const uint32_t rr = r < nrows ? r : 0;
        xrow[rb] = ((rr & 1U) ? ts_v : ts_e) + static_cast<uint64_t>(rr >> 1) * seglen;
    }

    // k-step st: samples k0 + 32 st + 8g .. + 7 of the lane's rows, as bf16
    // high and low parts (index rb * kSteps3 + st)
    auto load_x = [&](uint32_t k0, bf16x8* xh, bf16x8* xl) {
#pragma unroll
        for (int q = 0; q < MRB * kSteps3; ++q) {
            const int rb     = q / kSteps3;
            const int st     = q % kSteps3;
            const uint32_t r = row_base + rb * 16 + li;
            const uint32_t k = k0 + 32 * st + 8 * g;
            float v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                v[i] = 0.0F;
            }
            if (r < nrows) {
                if (vec_ok) {
                    const f32x4 a = *reinterpret_cast<const f32x4*>(xrow[rb] + k);
                    const f32x4 b = *reinterpret_cast<const f32x4*>(xrow[rb] + k + 4);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        v[i]     = a[i];
                        v[4 + i] = b[i];
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        v[i] = (k + i < seglen) ? xrow[rb][k + i] : 0.0F;
                    }
                }
            }
            bf16x2 h[4], l[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                split_bf16x2(f32x2{v[2 * i], v[2 * i + 1]}, h[i], l[i]);
            }
            xh[q] = __builtin_shufflevector(__builtin_shufflevector(h[0], h[1], 0, 1, 2, 3),
                                            __builtin_shufflevector(h[2], h[3], 0, 1, 2, 3),
                                            0, 1, 2, 3, 4, 5, 6, 7);
            xl[q] = __builtin_shufflevector(__builtin_shufflevector(l[0], l[1], 0, 1, 2, 3),
                                            __builtin_shufflevector(l[2], l[3], 0, 1, 2, 3),
                                            0, 1, 2, 3, 4, 5, 6, 7);
        }
    };

    // Twiddles of the chunk: column 2(m-1) = cos, 2(m-1)+1 = -sin of harmonic m.
    // Per sample, harmonics come in groups of kGrp: the first of a group is the
    // accurate twiddle of its FP64-reduced angle, the others follow by FP32
    // complex products with the harmonic-1 twiddle (a few ulp after kGrp - 1
    // steps, far below the bf16 split's 2^-17).
    constexpr uint32_t kGrp = 8;
    auto gen_b = [&](uint32_t k0, __bf16* th, __bf16* tl) {
        const uint32_t ngrp  = (nh + kGrp - 1) / kGrp;
        const uint32_t items = kKc3 * ngrp;
        for (uint32_t p = threadIdx.x; p < items; p += kDft3Block) {
            // consecutive lanes take consecutive samples: their 2-byte writes fill
            // whole LDS words of one column (no bank conflicts)
            const uint32_t kk = p % kKc3;
            const uint32_t m0 = (p / kKc3) * kGrp + 1;
            const uint32_t m1 = m0 + kGrp <= nh + 1 ? m0 + kGrp : nh + 1;
            const double ph   = f * ((static_cast<double>(k0 + kk) * tsamp) - t_ref);
            const double fr   = ph - floor(ph);
            const float2 w1   = twiddle_turns(fr);
            float2 w          = twiddle_turns(fr * static_cast<double>(m0));
            for (uint32_t m = m0; m < m1; ++m) {
                const uint32_t c0 = (2 * (m - 1)) * kKs3 + kk;
                bf16x2 h, l;
                split_bf16x2(f32x2{w.x, w.y}, h, l);
                th[c0]        = h[0];
                tl[c0]        = l[0];
                th[c0 + kKs3] = h[1];
                tl[c0 + kKs3] = l[1];
                w = make_float2(fmaf(w.x, w1.x, -(w.y * w1.y)), fmaf(w.x, w1.y, w.y * w1.x));
            }
        }
    };

    f32x4 acc[MRB][NCB], tot[MRB][NCB];
#pragma unroll
    for (int rb = 0; rb < MRB; ++rb) {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            acc[rb][cb] = f32x4{0.0F, 0.0F, 0.0F, 0.0F};
            tot[rb][cb] = f32x4{0.0F, 0.0F, 0.0F, 0.0F};
        }
    }

    const uint32_t nchunks = (seglen + kKc3 - 1) / kKc3;
    __syncthreads();  // zero-fill of the tables done
    gen_b(0, bt[0][0], bt[0][1]);
    bf16x8 xh[MRB * kSteps3], xl[MRB * kSteps3];
    load_x(0, xh, xl);
    __syncthreads();
    for (uint32_t ch = 0; ch < nchunks; ++ch) {
        const __bf16* th = bt[ch & 1U][0];
        const __bf16* tl = bt[ch & 1U][1];
        bf16x8 nxh[MRB * kSteps3], nxl[MRB * kSteps3];
        const bool more = ch + 1 < nchunks;
        if (more) {
            load_x((ch + 1) * kKc3, nxh, nxl);
            gen_b((ch + 1) * kKc3, bt[(ch + 1) & 1U][0], bt[(ch + 1) & 1U][1]);
        }
#pragma unroll
        for (int st = 0; st < kSteps3; ++st) {
            bf16x8 bh[NCB], bl[NCB];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const uint32_t o = (cb * 16 + li) * kKs3 + 32 * st + 8 * g;
                bh[cb] = *reinterpret_cast<const bf16x8*>(th + o);
                bl[cb] = *reinterpret_cast<const bf16x8*>(tl + o);
            }
#pragma unroll
            for (int rb = 0; rb < MRB; ++rb) {
                const int q = rb * kSteps3 + st;
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) {
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl[q], bh[cb],
                                                                          acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[q], bl[cb],
                                                                          acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[q], bh[cb],
                                                                          acc[rb][cb], 0, 0, 0);
                }
            }
        }
        if (more) {
#pragma unroll
            for (int q = 0; q < MRB * kSteps3; ++q) {
                xh[q] = nxh[q];
                xl[q] = nxl[q];
            }
        }
        if ((ch + 1) % kBlockChunks3 == 0 || !more) {
#pragma unroll
            for (int rb = 0; rb < MRB; ++rb) {
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) {
                    tot[rb][cb] += acc[rb][cb];
                    acc[rb][cb] = f32x4{0.0F, 0.0F, 0.0F, 0.0F};
                }
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int rb = 0; rb < MRB; ++rb) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint32_t r = row_base + rb * 16 + 4 * g + v;
            if (r >= nrows) {
                continue;
            }
            const uint64_t s = r >> 1;
            const uint32_t q = r & 1U;
            float* out = reinterpret_cast<float*>(fold + ((s * nfreqs + fi) * 2 + q) * nbf) + 2;
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const uint32_t n = cb * 16 + li;
                if (n < ncols) {
                    out[n] = tot[rb][cb][v];
                }
            }
        }
    }
}

template <int NCB>
void launch_bf_mfma3(const float* ts_e, const float* ts_v, float2* fold, const double* freqs,
                     uint32_t nfreqs, uint32_t nseg, uint32_t seglen, uint32_t nbf, double tsamp,
                     double t_ref, hipStream_t s) {
    constexpr uint32_t GROWS = 16 * dft_mrb(NCB) * kDft3Waves;
    const uint64_t nrb = (2ULL * nseg + GROWS - 1) / GROWS;
    const uint64_t nb  = nrb * nfreqs;
    LOKI_REQUIRE(nb < (1ULL << 31), "Fourier brute fold grid too large");
    k_bf_fourier_mfma3<NCB><<<static_cast<uint32_t>(nb), kDft3Block, 0, s>>>(
        ts_e, ts_v, fold, freqs, nfreqs, nseg, seglen, nbf, tsamp, t_ref);
}

template <int NCB>
void launch_bf_mfma(const float* ts_e, const float* ts_v, float2* fold, const double* freqs,
                    uint32_t nfreqs, uint32_t nseg, uint32_t seglen, uint32_t nbf, double tsamp,
                    double t_ref, hipStream_t s) {
    constexpr uint32_t GROWS = 16 * dft_mrb(NCB) * kDftWaves;
    const uint64_t nrb = (2ULL * nseg + GROWS - 1) / GROWS;
    const uint64_t nb  = nrb * nfreqs;
    LOKI_REQUIRE(nb < (1ULL << 31), "Fourier brute fold grid too large");
    k_bf_fourier_mfma<NCB><<<static_cast<uint32_t>(nb), kDftBlock, 0, s>>>(
        ts_e, ts_v, fold, freqs, nfreqs, nseg, seglen, nbf, tsamp, t_ref);
}

// ---------------------------------------------------------------------------
// Merge
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 w) {
    return make_float2(fmaf(a.x, w.x, -(a.y * w.y)), fmaf(a.x, w.y, a.y * w.x));
}

// IDX: uint32_t when nseg * ncur * nbf < 2^32 (cheaper index division), else uint64_t.
template <bool POLY, typename IDX>
__global__ void __launch_bounds__(kBlock)
k_merge_fourier(const float2* __restrict__ in, float2* __restrict__ out,
                const FFACoord* __restrict__ coords, const FFACoordFreq* __restrict__ cf,
                IDX ncur, uint64_t nprev, IDX nseg, uint32_t nbf, uint32_t nbins) {
    const IDX total    = nseg * ncur * nbf;
    const double inv_n = 1.0 / static_cast<double>(nbins);
    for (IDX t = static_cast<IDX>(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += static_cast<IDX>(gridDim.x) * kBlock) {
        const uint32_t k   = static_cast<uint32_t>(t % nbf);
        const IDX r        = t / nbf;
        const uint64_t c   = r % ncur;
        const uint64_t seg = r / ncur;
        uint64_t it_row, ih_row;
        float st = 0.0F, sh;
        if constexpr (POLY) {
            const FFACoord cd = coords[c];
            it_row = (2 * seg) * nprev + cd.i_tail;
            ih_row = (2 * seg + 1) * nprev + cd.i_head;
            st     = cd.shift_tail;
            sh     = cd.shift_head;
        } else {
            const FFACoordFreq cd = cf[c];
            it_row = (2 * seg) * nprev + cd.idx;
            ih_row = (2 * seg + 1) * nprev + cd.idx;
            sh     = cd.shift;
        }
        const float2* tr = in + it_row * 2 * nbf;
        const float2* hr = in + ih_row * 2 * nbf;
        const float2 te  = tr[k];
        const float2 tv  = tr[nbf + k];
        const float2 he  = hr[k];
        const float2 hv  = hr[nbf + k];
        const double kd  = static_cast<double>(k);
        const float2 wh  = twiddle_turns(kd * static_cast<double>(sh) * inv_n);
        float2 oe, ov;
        if constexpr (POLY) {
            const float2 wt = twiddle_turns(kd * static_cast<double>(st) * inv_n);
            const float2 a  = cmul(te, wt);
            const float2 b  = cmul(he, wh);
            const float2 cc = cmul(tv, wt);
            const float2 d  = cmul(hv, wh);
            oe = make_float2(a.x + b.x, a.y + b.y);
            ov = make_float2(cc.x + d.x, cc.y + d.y);
        } else {
            const float2 b = cmul(he, wh);
            const float2 d = cmul(hv, wh);
            oe = make_float2(te.x + b.x, te.y + b.y);
            ov = make_float2(tv.x + d.x, tv.y + d.y);
        }
        float2* o = out + (seg * ncur + c) * 2 * nbf;
        o[k]       = oe;
        o[nbf + k] = ov;
    }
}

// ---------------------------------------------------------------------------
// Real transforms along rows of nbins (fft_cuda.cu:282-448: cuFFT R2C, and
// C2R followed by the 1/n scale)
// ---------------------------------------------------------------------------
constexpr uint32_t kMaxTwiddle = 4096;

// x_j = (1/n) [X_0 + 2 sum_{k=1}^{kmax-1} Re(X_k e^{2 pi i jk/n}) + X_{n/2} (-1)^j (n even)]
__global__ void __launch_bounds__(kBlock)
k_irfft(const float2* __restrict__ in, float* __restrict__ out, uint64_t nrows, uint32_t nbins,
        float norm) {
    extern __shared__ float2 tw[];  // exp(2 pi i t / n), t in [0, n)
    for (uint32_t t = threadIdx.x; t < nbins; t += kBlock) {
        const double a = static_cast<double>(t) / static_cast<double>(nbins);
        const float2 w = twiddle_turns(a);
        tw[t]          = make_float2(w.x, -w.y);
    }
    __syncthreads();
    const uint32_t nbf  = nbins / 2 + 1;
    const uint32_t kmax = (nbins % 2 == 0) ? nbins / 2 : nbins / 2 + 1;
    const uint64_t total = nrows * nbins;
    for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint32_t j  = static_cast<uint32_t>(t % nbins);
        const uint64_t r  = t / nbins;
        const float2* X   = in + r * nbf;
        float acc         = 0.0F;
        uint32_t idx      = j;
        for (uint32_t k = 1; k < kmax; ++k) {
            const float2 x = X[k];
            const float2 w = tw[idx];
            acc            = fmaf(x.x, w.x, acc);
            acc            = fmaf(-x.y, w.y, acc);
            idx += j;
            idx = idx >= nbins ? idx - nbins : idx;
        }
        float sum = fmaf(2.0F, acc, X[0].x);
        if (nbins % 2 == 0) {
            const float xn = X[nbins / 2].x;
            sum            = (j & 1U) ? sum - xn : sum + xn;
        }
        out[t] = sum * norm;
    }
}

// X_k = sum_j x_j e^{-2 pi i jk/n}, k = 0..n/2 (unnormalised)
__global__ void __launch_bounds__(kBlock)
k_rfft(const float* __restrict__ in, float2* __restrict__ out, uint64_t nrows, uint32_t nbins) {
    extern __shared__ float2 tw[];  // exp(-2 pi i t / n)
    for (uint32_t t = threadIdx.x; t < nbins; t += kBlock) {
        tw[t] = twiddle_turns(static_cast<double>(t) / static_cast<double>(nbins));
    }
    __syncthreads();
    const uint32_t nbf   = nbins / 2 + 1;
    const uint64_t total = nrows * nbf;
    for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint32_t k = static_cast<uint32_t>(t % nbf);
        const uint64_t r = t / nbf;
        const float* x   = in + r * nbins;
        float re = 0.0F, im = 0.0F;
        uint32_t idx = 0;
        for (uint32_t j = 0; j < nbins; ++j) {
            const float2 w = tw[idx];
            re             = fmaf(x[j], w.x, re);
            im             = fmaf(x[j], w.y, im);
            idx += k;
            idx = idx >= nbins ? idx - nbins : idx;
        }
        out[t] = make_float2(re, im);
    }
}

unsigned grid_for(uint64_t total) {
    const uint64_t want = (total + kBlock - 1) / kBlock;
    const uint64_t cap  = 65536ULL * 8;
    return static_cast<unsigned>(want < cap ? (want == 0 ? 1 : want) : cap);
}

}  // namespace

void brute_fold_fourier(const float* ts_e, const float* ts_v, uint64_t nseg, uint32_t seglen,
                        const double* freqs_dev, uint32_t nfreqs, uint32_t nbins, double tsamp,
                        double t_ref, float2* fold, hipStream_t s, bool split_bf16) {
    LOKI_REQUIRE(nbins >= 2 && nbins <= 256, "Fourier brute fold: nbins must be in [2, 256]");
    LOKI_REQUIRE(seglen >= 1 && nfreqs >= 1, "Fourier brute fold: empty input");
    LOKI_REQUIRE(nseg >= 1 && 2 * nseg < (1ULL << 31), "Fourier brute fold: too many segments");
    const uint32_t nbf = nbins / 2 + 1;
    k_bf_fourier_dc<<<static_cast<uint32_t>(nseg), kBlock, 0, s>>>(ts_e, ts_v, fold, seglen,
                                                                   nfreqs, nbf);
    const uint32_t ncols = 2 * (nbf - 1);
    const auto ns        = static_cast<uint32_t>(nseg);
    if (split_bf16 && ncols <= 128) {
        if (ncols <= 16) {
            launch_bf_mfma3<1>(ts_e, ts_v, fold, freqs_dev, nfreqs, ns, seglen, nbf, tsamp, t_ref,
                               s);
        } else if (ncols <= 32) {
            launch_bf_mfma3<2>(ts_e, ts_v, fold, freqs_dev, nfreqs, ns, seglen, nbf, tsamp, t_ref,
                               s);
        } else if (ncols <= 64) {
            launch_bf_mfma3<4>(ts_e, ts_v, fold, freqs_dev, nfreqs, ns, seglen, nbf, tsamp, t_ref,
                               s);
        } else {
            launch_bf_mfma3<8>(ts_e, ts_v, fold, freqs_dev, nfreqs, ns, seglen, nbf, tsamp, t_ref,
                               s);
        }
        return;
    }
    if (ncols <= 16) {
        launch_bf_mfma<1>(ts_e, ts_v, fold, freqs_dev, nfreqs, ns, seglen, nbf, tsamp, t_ref, s);
    } else if (ncols <= 32) {
        launch_bf_mfma<2>(ts_e, ts_v, fold, freqs_dev, nfreqs, ns, seglen, nbf, tsamp, t_ref, s);
    } else if (ncols <= 64) {
        launch_bf_mfma<4>(ts_e, ts_v, fold, freqs_dev, nfreqs, ns, seglen, nbf, tsamp, t_ref, s);
    } else if (ncols <= 128) {
        launch_bf_mfma<8>(ts_e, ts_v, fold, freqs_dev, nfreqs, ns, seglen, nbf, tsamp, t_ref, s);
    } else {
        launch_bf_mfma<16>(ts_e, ts_v, fold, freqs_dev, nfreqs, ns, seglen, nbf, tsamp, t_ref,
                           s);
    }
}

void merge_fourier(const float2* in, float2* out, const FFACoord* coords,
                   const FFACoordFreq* coords_freq, uint64_t ncur, uint64_t nprev, uint64_t nseg,
                   uint32_t nbins, hipStream_t s) {
    LOKI_REQUIRE(nbins >= 1, "merge: nbins must be positive");
    LOKI_REQUIRE((coords == nullptr) != (coords_freq == nullptr),
                 "merge: exactly one coordinate table");
    if (ncur == 0 || nseg == 0) {
        return;
    }
    const uint32_t nbf   = nbins / 2 + 1;
    const uint64_t total = nseg * ncur * nbf;
    const unsigned g     = grid_for(total);
    // the grid-stride loop must not overflow the index type
    const bool narrow = total + static_cast<uint64_t>(g) * kBlock < (1ULL << 32);
    if (coords != nullptr) {
        if (narrow) {
            k_merge_fourier<true, uint32_t><<<g, kBlock, 0, s>>>(
                in, out, coords, nullptr, static_cast<uint32_t>(ncur), nprev,
                static_cast<uint32_t>(nseg), nbf, nbins);
        } else {
            k_merge_fourier<true, uint64_t><<<g, kBlock, 0, s>>>(in, out, coords, nullptr, ncur,
                                                                 nprev, nseg, nbf, nbins);
        }
    } else {
        if (narrow) {
            k_merge_fourier<false, uint32_t><<<g, kBlock, 0, s>>>(
                in, out, nullptr, coords_freq, static_cast<uint32_t>(ncur), nprev,
                static_cast<uint32_t>(nseg), nbf, nbins);
        } else {
            k_merge_fourier<false, uint64_t><<<g, kBlock, 0, s>>>(
                in, out, nullptr, coords_freq, ncur, nprev, nseg, nbf, nbins);
        }
    }
}

void irfft_rows(const float2* in, float* out, uint64_t nrows, uint32_t nbins, hipStream_t s) {
    LOKI_REQUIRE(nbins >= 1 && nbins <= kMaxTwiddle, "irfft: nbins must be in [1, 4096]");
    if (nrows == 0) {
        return;
    }
    const float norm = 1.0F / static_cast<float>(nbins);
    k_irfft<<<grid_for(nrows * nbins), kBlock, nbins * sizeof(float2), s>>>(in, out, nrows, nbins,
                                                                            norm);
}

void rfft_rows(const float* in, float2* out, uint64_t nrows, uint32_t nbins, hipStream_t s) {
    LOKI_REQUIRE(nbins >= 1 && nbins <= kMaxTwiddle, "rfft: nbins must be in [1, 4096]");
    if (nrows == 0) {
        return;
    }
    k_rfft<<<grid_for(nrows * (nbins / 2 + 1)), kBlock, nbins * sizeof(float2), s>>>(in, out,
                                                                                     nrows, nbins);
}

}  // namespace loki_amd::kern
