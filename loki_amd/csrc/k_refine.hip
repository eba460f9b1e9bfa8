// Differentiable Fourier-domain candidate score and its adjoint.
//
// For candidates theta_c = (f, fdot, fddot) the whole series is folded in the
// Fourier domain at the continuous ephemeris
//   phi(t) = f tau + fdot tau^2 / 2 + fddot tau^3 / 6,  tau = k tsamp - t_ref
// (the Fourier brute fold of kernels_cuda.cu:780-853 over one segment spanning
// the series, with the Taylor phase model of transforms.cpp:101-120):
//   E_m = sum_k e_k w_mk,  V_m = sum_k v_k w_mk,  w_mk = exp(-2 pi i m frac(phi_k)),
// then returned to nbins phase bins (1/n inverse real DFT, fft_cuda.cu:357-448),
// normalised p_j = e_j / sqrt(v_j) and scored with the fixed-width boxcar of
// score.cpp:95-172 at its best position.  The adjoint holds that position fixed:
//   dS/dphi_k = sum_m 2 pi m (a_m Im w_mk - b_m Re w_mk) e_k + (same with V) v_k
// with (a_m + i b_m) = dS/dRe E_m + i dS/dIm E_m, and dS/dtheta = sum_k dS/dphi_k
// (tau, tau^2/2, tau^3/6).  Nothing but the per-candidate harmonics is stored:
// the fold is recomputed from the series for the adjoint pass.
//
// Phases are FP64; w_mk comes from an accurate sincospi of the FP64-reduced
// phase and an FP32 three-term power recurrence; per-lane partial sums are
// FP32, block and grid reductions FP64.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"

namespace loki_amd::kern {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ float2 turns_twiddle(double d) {
    d -= floor(d);
    if (d > 0.5) {
        d -= 1.0;
    }
    const float hi = static_cast<float>(d);
    const float lo = static_cast<float>(d - static_cast<double>(hi));
    float sn, cs;
    sincospif(2.0F * hi, &sn, &cs);
    const float dl = 6.28318530717958647692F * lo;
    return make_float2(fmaf(-dl, sn, cs), -fmaf(dl, cs, sn));  // exp(-2 pi i d)
}

__device__ __forceinline__ double phase_at(uint64_t k, double tsamp, double t_ref, double f,
                                           double fd, double fdd, double* tau_out) {
    const double tau = (static_cast<double>(k) * tsamp) - t_ref;
    *tau_out         = tau;
    return (f * tau) + ((fd * tau * tau) / 2.0) + ((fdd * tau * tau * tau) / 6.0);
}

/// Sum of `x` over the 256 lanes of the block (FP64), valid in thread 0.
__device__ __forceinline__ double block_sum(double x, double* red) {
    for (int o = 32; o > 0; o >>= 1) {
        x += __shfl_xor(x, o);
    }
    __syncthreads();
    if ((threadIdx.x & 63U) == 0) {
        red[threadIdx.x >> 6] = x;
    }
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < kWaves; ++i) {
            s += red[i];
        }
    }
    return s;
}

/// Sum of `x` over the 64 lanes of the wave (FP32), valid in every lane.
__device__ __forceinline__ float wave_sum(float x) {
    for (int o = 32; o > 0; o >>= 1) {
        x += __shfl_xor(x, o);
    }
    return x;
}

/// Powers w^m of a unit twiddle w = (cos t, -sin t) by the three-term
/// (Chebyshev) recurrence w^{m+1} = 2 cos(t) w^m - w^{m-1}, applied to the real
/// and imaginary parts separately: two FMAs per power instead of a complex
/// product.  Its rounding error grows at most like m^2/2 ulps (a perturbation
/// at step j reaches step m scaled by |sin((m-j)t)/sin(t)| <= m - j), ~3e-5 at
/// m = 32, against ~m ulps for repeated complex products.
struct Powers {
    float pr, pi;  // w^m (the power in use)
    float cr, ci;  // w^{m+1}
    float c2;      // 2 cos t
    __device__ __forceinline__ explicit Powers(float2 z)
        : pr(1.0F), pi(0.0F), cr(z.x), ci(z.y), c2(2.0F * z.x) {}
    __device__ __forceinline__ void step() {
        const float nr = fmaf(c2, cr, -pr);
        const float ni = fmaf(c2, ci, -pi);
        pr = cr;
        pi = ci;
        cr = nr;
        ci = ni;
    }
};

// Forward partials: block (chunk q, candidate c) -> part[c][q][NH][4] =
// (Re E, Im E, Re V, Im V) of its samples.  Each lane takes two samples per
// iteration (two independent power recurrences); lane sums are FP32, the
// wave partials are added in FP64.
template <int NH>
__global__ void __launch_bounds__(kBlock)
k_refine_fold(const float* __restrict__ ts_e, const float* __restrict__ ts_v, uint64_t nsamps,
              uint64_t chunk, double tsamp, double t_ref, const double* __restrict__ theta,
              double* __restrict__ part) {
    __shared__ float red[kWaves][4 * NH];
    const uint32_t c  = blockIdx.y;
    const uint64_t k0 = static_cast<uint64_t>(blockIdx.x) * chunk;
    const uint64_t k1 = k0 + chunk < nsamps ? k0 + chunk : nsamps;
    const double f = theta[3 * c], fd = theta[3 * c + 1], fdd = theta[3 * c + 2];
    float er[NH], ei[NH], vr[NH], vi[NH];
#pragma unroll
    for (int m = 0; m < NH; ++m) {
        er[m] = ei[m] = vr[m] = vi[m] = 0.0F;
    }
    for (uint64_t k = k0 + threadIdx.x; k < k1; k += 2 * kBlock) {
        const uint64_t kb = k + kBlock;
        const bool has_b  = kb < k1;
        double tau;
        const float2 za = turns_twiddle(phase_at(k, tsamp, t_ref, f, fd, fdd, &tau));
        const float2 zb = turns_twiddle(phase_at(has_b ? kb : k, tsamp, t_ref, f, fd, fdd, &tau));
        const float xa = ts_e[k], ya = ts_v[k];
        const float xb = has_b ? ts_e[kb] : 0.0F, yb = has_b ? ts_v[kb] : 0.0F;
        Powers wa(za), wb(zb);
#pragma unroll
        for (int m = 0; m < NH; ++m) {
            er[m] = fmaf(xb, wb.pr, fmaf(xa, wa.pr, er[m]));
            ei[m] = fmaf(xb, wb.pi, fmaf(xa, wa.pi, ei[m]));
            vr[m] = fmaf(yb, wb.pr, fmaf(ya, wa.pr, vr[m]));
            vi[m] = fmaf(yb, wb.pi, fmaf(ya, wa.pi, vi[m]));
            wa.step();
            wb.step();
        }
    }
    const uint32_t lane = threadIdx.x & 63U, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < NH; ++m) {
        const float s0 = wave_sum(er[m]);
        const float s1 = wave_sum(ei[m]);
        const float s2 = wave_sum(vr[m]);
        const float s3 = wave_sum(vi[m]);
        if (lane == 0) {
            red[wave][4 * m + 0] = s0;
            red[wave][4 * m + 1] = s1;
            red[wave][4 * m + 2] = s2;
            red[wave][4 * m + 3] = s3;
        }
    }
    __syncthreads();
    double* out = part + (static_cast<uint64_t>(c) * gridDim.x + blockIdx.x) * (4 * NH);
    for (uint32_t i = threadIdx.x; i < 4 * NH; i += kBlock) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            s += static_cast<double>(red[w][i]);
        }
        out[i] = s;
    }
}

// Per candidate (one 64-lane block): harmonics -> profile -> S/N, and the
// adjoint vectors state[c] = (dS/dE_m, dS/dV_m) as complex (FP32).
template <int NH>
__global__ void __launch_bounds__(64)
k_refine_score(const double* __restrict__ part, uint32_t nchunk, uint32_t nbins, uint32_t width,
               float* __restrict__ scores, float2* __restrict__ state) {
    constexpr int NBMAX = 2 * (NH - 1);
    __shared__ double s_h[4 * NH];
    __shared__ double s_e[NBMAX], s_v[NBMAX], s_p[NBMAX], s_de[NBMAX], s_dv[NBMAX];
    __shared__ int s_j;
    const uint32_t c = blockIdx.x;
    const double* pc = part + static_cast<uint64_t>(c) * nchunk * (4 * NH);
    for (uint32_t i = threadIdx.x; i < 4 * NH; i += 64) {
        double s = 0.0;
        for (uint32_t q = 0; q < nchunk; ++q) {
            s += pc[static_cast<uint64_t>(q) * (4 * NH) + i];
        }
        s_h[i] = s;
    }
    __syncthreads();
    const uint32_t n  = nbins;
    const uint32_t nh = n / 2 + 1;
    const double two_pi_n = 6.283185307179586476925 / static_cast<double>(n);
    // inverse real DFT (1/n), imaginary parts of X_0 and X_{n/2} ignored
    for (uint32_t j = threadIdx.x; j < n; j += 64) {
        double ae = s_h[0], av = s_h[2];
        for (uint32_t m = 1; m + 1 < nh; ++m) {
            const double a  = two_pi_n * static_cast<double>((static_cast<uint64_t>(j) * m) % n);
            const double cs = cos(a), sn = sin(a);
            ae += 2.0 * (s_h[4 * m] * cs - s_h[4 * m + 1] * sn);
            av += 2.0 * (s_h[4 * m + 2] * cs - s_h[4 * m + 3] * sn);
        }
        const double sg = (j & 1U) ? -1.0 : 1.0;
        ae += s_h[4 * (nh - 1)] * sg;
        av += s_h[4 * (nh - 1) + 2] * sg;
        s_e[j] = ae / static_cast<double>(n);
        s_v[j] = av / static_cast<double>(n);
        s_p[j] = s_e[j] / sqrt(s_v[j]);
    }
    __syncthreads();
    const double w  = static_cast<double>(width);
    const double hh = sqrt((static_cast<double>(n) - w) / (static_cast<double>(n) * w));
    const double bb = w * hh / (static_cast<double>(n) - w);
    if (threadIdx.x == 0) {
        double best = -1.0e300, sum = 0.0;
        int jb = 0;
        for (uint32_t j = 0; j < n; ++j) {
            sum += s_p[j];
            double wsum = 0.0;
            for (uint32_t i = 0; i < width; ++i) {
                wsum += s_p[(j + i) % n];
            }
            if (wsum > best) {
                best = wsum;
                jb   = static_cast<int>(j);
            }
        }
        scores[c] = static_cast<float>((hh + bb) * best - bb * sum);
        s_j       = jb;
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n; j += 64) {
        const uint32_t off = (j + n - static_cast<uint32_t>(s_j)) % n;
        const double dp    = (off < width ? hh + bb : 0.0) - bb;
        const double rv    = 1.0 / sqrt(s_v[j]);
        s_de[j] = dp * rv;
        s_dv[j] = -dp * s_e[j] * rv / (2.0 * s_v[j]);
    }
    __syncthreads();
    // dS/dRe X_m + i dS/dIm X_m through the inverse DFT
    if (state != nullptr) {
        for (uint32_t m = threadIdx.x; m < nh; m += 64) {
            double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
            const double cm = (m == 0 || m == nh - 1) ? 1.0 : 2.0;
            for (uint32_t j = 0; j < n; ++j) {
                const double a  = two_pi_n * static_cast<double>((static_cast<uint64_t>(j) * m) % n);
                const double cs = cos(a), sn = sin(a);
                ar += s_de[j] * cs;
                ai += -s_de[j] * sn;
                br += s_dv[j] * cs;
                bi += -s_dv[j] * sn;
            }
            const double k  = cm / static_cast<double>(n);
            const bool real_only = (m == 0 || m == nh - 1);
            float2* st = state + static_cast<uint64_t>(c) * (2 * nh);
            st[m]      = make_float2(static_cast<float>(k * ar),
                                     real_only ? 0.0F : static_cast<float>(k * ai));
            st[nh + m] = make_float2(static_cast<float>(k * br),
                                     real_only ? 0.0F : static_cast<float>(k * bi));
        }
    }
}

// Adjoint partials: block (chunk q, candidate c) -> gpart[c][q][3].
template <int NH>
__global__ void __launch_bounds__(kBlock)
k_refine_adjoint(const float* __restrict__ ts_e, const float* __restrict__ ts_v, uint64_t nsamps,
                 uint64_t chunk, double tsamp, double t_ref, const double* __restrict__ theta,
                 const float2* __restrict__ state, double* __restrict__ gpart) {
    __shared__ double red[kWaves];
    __shared__ float2 s_st[2 * NH];
    const uint32_t c = blockIdx.y;
    for (uint32_t i = threadIdx.x; i < 2 * NH; i += kBlock) {
        s_st[i] = state[static_cast<uint64_t>(c) * (2 * NH) + i];
    }
    __syncthreads();
    float ge[NH], gi_e[NH], gv[NH], gi_v[NH];  // m * (a, b) of E and V
#pragma unroll
    for (int m = 0; m < NH; ++m) {
        ge[m]   = static_cast<float>(m) * s_st[m].x;
        gi_e[m] = static_cast<float>(m) * s_st[m].y;
        gv[m]   = static_cast<float>(m) * s_st[NH + m].x;
        gi_v[m] = static_cast<float>(m) * s_st[NH + m].y;
    }
    const uint64_t k0 = static_cast<uint64_t>(blockIdx.x) * chunk;
    const uint64_t k1 = k0 + chunk < nsamps ? k0 + chunk : nsamps;
    const double f = theta[3 * c], fd = theta[3 * c + 1], fdd = theta[3 * c + 2];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (uint64_t k = k0 + threadIdx.x; k < k1; k += 2 * kBlock) {
        const uint64_t kb = k + kBlock;
        const bool has_b  = kb < k1;
        double ta, tb;
        const float2 za = turns_twiddle(phase_at(k, tsamp, t_ref, f, fd, fdd, &ta));
        const float2 zb = turns_twiddle(phase_at(has_b ? kb : k, tsamp, t_ref, f, fd, fdd, &tb));
        Powers wa(za), wb(zb);
        float dea = 0.0F, dva = 0.0F, deb = 0.0F, dvb = 0.0F;
#pragma unroll
        for (int m = 0; m < NH; ++m) {
            dea = fmaf(-gi_e[m], wa.pr, fmaf(ge[m], wa.pi, dea));
            dva = fmaf(-gi_v[m], wa.pr, fmaf(gv[m], wa.pi, dva));
            deb = fmaf(-gi_e[m], wb.pr, fmaf(ge[m], wb.pi, deb));
            dvb = fmaf(-gi_v[m], wb.pr, fmaf(gv[m], wb.pi, dvb));
            wa.step();
            wb.step();
        }
        const double pa = static_cast<double>(dea) * ts_e[k] + static_cast<double>(dva) * ts_v[k];
        const double pb = has_b ? static_cast<double>(deb) * ts_e[kb] +
                                      static_cast<double>(dvb) * ts_v[kb]
                                : 0.0;
        const double da = 6.283185307179586476925 * pa;
        const double db = 6.283185307179586476925 * pb;
        g0 += da * ta + db * tb;
        g1 += (da * ta * ta + db * tb * tb) / 2.0;
        g2 += (da * ta * ta * ta + db * tb * tb * tb) / 6.0;
    }
    const double s0 = block_sum(g0, red);
    const double s1 = block_sum(g1, red);
    const double s2 = block_sum(g2, red);
    if (threadIdx.x == 0) {
        double* out = gpart + (static_cast<uint64_t>(c) * gridDim.x + blockIdx.x) * 3;
        out[0]      = s0;
        out[1]      = s1;
        out[2]      = s2;
    }
}

__global__ void __launch_bounds__(64)
k_refine_grad_sum(const double* __restrict__ gpart, uint32_t nchunk, uint32_t ncand,
                  double* __restrict__ grad) {
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= 3 * ncand) {
        return;
    }
    const uint32_t c = i / 3, d = i % 3;
    double s = 0.0;
    for (uint32_t q = 0; q < nchunk; ++q) {
        s += gpart[(static_cast<uint64_t>(c) * nchunk + q) * 3 + d];
    }
    grad[i] = s;
}

uint32_t chunks_for(uint64_t nsamps, uint64_t ncand) {
    // about 8192 blocks in flight, chunks of >= 1024 samples
    ncand        = ncand < 1 ? 1 : ncand;
    uint64_t nq  = (8192 + ncand - 1) / ncand;
    const uint64_t max_q = (nsamps + 1023) / 1024;
    nq = nq < max_q ? nq : max_q;
    return static_cast<uint32_t>(nq < 1 ? 1 : nq);
}

template <int NH>
void forward_nh(const float* ts_e, const float* ts_v, uint64_t nsamps, double tsamp,
                double t_ref, const double* theta, uint32_t ncand, uint32_t nbins,
                uint32_t width, float* scores, float2* state, double* part, uint32_t nq,
                hipStream_t s) {
    const uint64_t chunk = (nsamps + nq - 1) / nq;
    k_refine_fold<NH><<<dim3(nq, ncand), kBlock, 0, s>>>(ts_e, ts_v, nsamps, chunk, tsamp, t_ref,
                                                         theta, part);
    k_refine_score<NH><<<ncand, 64, 0, s>>>(part, nq, nbins, width, scores, state);
}

template <int NH>
void adjoint_nh(const float* ts_e, const float* ts_v, uint64_t nsamps, double tsamp,
                double t_ref, const double* theta, uint32_t ncand, const float2* state,
                double* gpart, double* grad, uint32_t nq, hipStream_t s) {
    const uint64_t chunk = (nsamps + nq - 1) / nq;
    k_refine_adjoint<NH><<<dim3(nq, ncand), kBlock, 0, s>>>(ts_e, ts_v, nsamps, chunk, tsamp,
                                                            t_ref, theta, state, gpart);
    k_refine_grad_sum<<<(3 * ncand + 63) / 64, 64, 0, s>>>(gpart, nq, ncand, grad);
}

}  // namespace

SizeType refine_workspace_bytes(uint64_t nsamps, uint64_t ncand, uint32_t nbins) {
    const uint64_t nq = chunks_for(nsamps, ncand);
    const uint64_t nh = nbins / 2 + 1;
    return ncand * nq * (4 * nh) * sizeof(double);
}

void refine_forward(const float* ts_e, const float* ts_v, uint64_t nsamps, double tsamp,
                    double t_ref, const double* theta, uint64_t ncand, uint32_t nbins,
                    uint32_t width, float* scores, float2* state, void* workspace,
                    hipStream_t s) {
    LOKI_REQUIRE(nbins == 16 || nbins == 32 || nbins == 64 || nbins == 128,
                 "refine: nbins must be 16, 32, 64 or 128");
    LOKI_REQUIRE(width >= 1 && width < nbins, "refine: width must be in [1, nbins)");
    LOKI_REQUIRE(ncand >= 1 && ncand < 65536, "refine: 1 <= candidates < 65536");
    LOKI_REQUIRE(nsamps >= 1, "refine: empty series");
    const auto nq = chunks_for(nsamps, ncand);
    auto* part    = static_cast<double*>(workspace);
    const auto nc = static_cast<uint32_t>(ncand);
    switch (nbins) {
    case 16: return forward_nh<9>(ts_e, ts_v, nsamps, tsamp, t_ref, theta, nc, nbins, width, scores, state, part, nq, s);
    case 32: return forward_nh<17>(ts_e, ts_v, nsamps, tsamp, t_ref, theta, nc, nbins, width, scores, state, part, nq, s);
    case 64: return forward_nh<33>(ts_e, ts_v, nsamps, tsamp, t_ref, theta, nc, nbins, width, scores, state, part, nq, s);
    default: return forward_nh<65>(ts_e, ts_v, nsamps, tsamp, t_ref, theta, nc, nbins, width, scores, state, part, nq, s);
    }
}

void refine_adjoint(const float* ts_e, const float* ts_v, uint64_t nsamps, double tsamp,
                    double t_ref, const double* theta, uint64_t ncand, uint32_t nbins,
                    const float2* state, double* grad, void* workspace, hipStream_t s) {
    LOKI_REQUIRE(nbins == 16 || nbins == 32 || nbins == 64 || nbins == 128,
                 "refine: nbins must be 16, 32, 64 or 128");
    LOKI_REQUIRE(ncand >= 1 && ncand < 65536, "refine: 1 <= candidates < 65536");
    const auto nq = chunks_for(nsamps, ncand);
    auto* gpart   = static_cast<double*>(workspace);
    const auto nc = static_cast<uint32_t>(ncand);
    switch (nbins) {
    case 16: return adjoint_nh<9>(ts_e, ts_v, nsamps, tsamp, t_ref, theta, nc, state, gpart, grad, nq, s);
    case 32: return adjoint_nh<17>(ts_e, ts_v, nsamps, tsamp, t_ref, theta, nc, state, gpart, grad, nq, s);
    case 64: return adjoint_nh<33>(ts_e, ts_v, nsamps, tsamp, t_ref, theta, nc, state, gpart, grad, nq, s);
    default: return adjoint_nh<65>(ts_e, ts_v, nsamps, tsamp, t_ref, theta, nc, state, gpart, grad, nq, s);
    }
}

}  // namespace loki_amd::kern
