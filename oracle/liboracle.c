/*
 * liboracle.c -- CPU oracle for the FFA hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library; it is the checker and the timed CPU baseline, never the
 * product path.
 *
 * A plain-C (OpenMP) restatement of pravirkr/loki v0.0.3's CPU algorithm.
 * Built with -O2 -ffp-contract=off and without -ffast-math, so every float
 * operation is one IEEE-754 rounding in source order.  That order is the
 * definition of "bit-exact" for the time-domain path:
 *   - brute fold: per (freq, bin) bucket, FP32 sum starting from 0.0f in
 *     ascending sample order (kernels.cpp:550-580, fold.cpp:113-142);
 *   - merge: one FP32 add per output element, tail + head
 *     (kernels.cpp:71-146, 722-940);
 *   - boxcar S/N: p = e * (1/sqrt(v)), sequential circular prefix sum,
 *     max window difference, snr = (h+b)*dmax - b*sum (score.cpp:95-172,
 *     utils.cpp:22-67).
 * The Fourier-domain functions are FP64 textbook versions (exact twiddles),
 * the reference the GPU's FP32 Fourier path is held to within a tolerance.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORC_PI 3.14159265358979323846

typedef struct {
    uint32_t i_tail;
    float shift_tail;
    uint32_t i_head;
    float shift_head;
} orc_coord;

typedef struct {
    uint32_t idx;
    float shift;
} orc_coord_freq;

static int clamp_threads(int n) {
    int m = omp_get_max_threads();
    if (n < 1) return 1;
    return n > m ? m : n;
}

/* psr_utils.hpp:32-49 get_phase_idx (FP64, result rounded to float). */
float orc_phase_idx(double proper_time, double freq, size_t nbins, double delay) {
    const double total = (proper_time - delay) * freq;
    const double norm = total - floor(total);
    double iphase = norm * (double)nbins;
    if (iphase >= (double)nbins) iphase = 0.0;
    return (float)iphase;
}

/* psr_utils.hpp:53-63 get_phase_idx_uint: u32(f32 phase + 0.5f), nbins -> 0. */
uint32_t orc_phase_bin(double proper_time, double freq, size_t nbins) {
    const float ph = orc_phase_idx(proper_time, freq, nbins, 0.0);
    const float r = ph + 0.5f;
    uint32_t b = (uint32_t)r;
    if (b == nbins) b = 0;
    return b;
}

/* kernels.cpp:78-85: integer bin shift from a float phase shift. */
static inline size_t shift_bins(float s, size_t nbins) {
    const float r = s + 0.5f;
    size_t k = (size_t)r;
    if (k == nbins) k = 0;
    return k;
}

/*
 * Time-domain brute fold, fold.cpp:113-142 (bucket build) and
 * kernels.cpp:550-580 / 1132-1154 (bucket sums).  fold is
 * [nseg, nfreqs, 2, nbins]; each bucket is summed in ascending sample order.
 */
void orc_brute_fold_time(const float* ts_e, const float* ts_v, size_t nsamps,
                         const double* freqs, size_t nfreqs, size_t seglen,
                         size_t nbins, double tsamp, double t_ref, float* fold,
                         int nthreads) {
    const size_t nseg = nsamps / seglen;
    const size_t nbuck = nfreqs * nbins;
    size_t* offsets = (size_t*)calloc(nbuck + 1, sizeof(size_t));
    uint32_t* bins = (uint32_t*)malloc(nfreqs * seglen * sizeof(uint32_t));
    uint32_t* idx = (uint32_t*)malloc(nfreqs * seglen * sizeof(uint32_t));
    for (size_t f = 0; f < nfreqs; ++f) {
        for (size_t k = 0; k < seglen; ++k) {
            const double t = ((double)k * tsamp) - t_ref;
            const uint32_t b = orc_phase_bin(t, freqs[f], nbins);
            bins[f * seglen + k] = b;
            offsets[f * nbins + b + 1]++;
        }
    }
    for (size_t i = 1; i <= nbuck; ++i) offsets[i] += offsets[i - 1];
    size_t* wr = (size_t*)malloc(nbuck * sizeof(size_t));
    memcpy(wr, offsets, nbuck * sizeof(size_t));
    for (size_t f = 0; f < nfreqs; ++f)
        for (size_t k = 0; k < seglen; ++k)
            idx[wr[f * nbins + bins[f * seglen + k]]++] = (uint32_t)k;
    free(wr);
    free(bins);
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t s = 0; s < nseg; ++s) {
        const float* e = ts_e + s * seglen;
        const float* v = ts_v + s * seglen;
        float* out = fold + s * nfreqs * 2 * nbins;
        for (size_t f = 0; f < nfreqs; ++f) {
            for (size_t b = 0; b < nbins; ++b) {
                const size_t bk = f * nbins + b;
                float se = 0.0f, sv = 0.0f;
                for (size_t i = offsets[bk]; i < offsets[bk + 1]; ++i) {
                    se += e[idx[i]];
                    sv += v[idx[i]];
                }
                out[f * 2 * nbins + b] = se;
                out[f * 2 * nbins + nbins + b] = sv;
            }
        }
    }
    free(offsets);
    free(idx);
}

/* Same brute fold in FP64 (the FP64 oracle fold used for precision studies). */
void orc_brute_fold_time_f64(const float* ts_e, const float* ts_v, size_t nsamps,
                             const double* freqs, size_t nfreqs, size_t seglen,
                             size_t nbins, double tsamp, double t_ref,
                             double* fold, int nthreads) {
    const size_t nseg = nsamps / seglen;
    uint32_t* bins = (uint32_t*)malloc(nfreqs * seglen * sizeof(uint32_t));
    for (size_t f = 0; f < nfreqs; ++f)
        for (size_t k = 0; k < seglen; ++k)
            bins[f * seglen + k] = orc_phase_bin(((double)k * tsamp) - t_ref, freqs[f], nbins);
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t s = 0; s < nseg; ++s) {
        double* out = fold + s * nfreqs * 2 * nbins;
        memset(out, 0, nfreqs * 2 * nbins * sizeof(double));
        for (size_t f = 0; f < nfreqs; ++f) {
            const uint32_t* bf = bins + f * seglen;
            for (size_t k = 0; k < seglen; ++k) {
                out[f * 2 * nbins + bf[k]] += (double)ts_e[s * seglen + k];
                out[f * 2 * nbins + nbins + bf[k]] += (double)ts_v[s * seglen + k];
            }
        }
    }
    free(bins);
}

/*
 * o[j] = t[(j - St) mod n] + h[(j - Sh) mod n] for j < n, one FP32 add per
 * element in the order of kernels.cpp:71-146; written as up to three
 * contiguous runs (split at St and Sh) so that the compiler vectorises them.
 */
static inline void add_rot_f32(float* restrict o, const float* restrict t, size_t St,
                               const float* restrict h, size_t Sh, size_t n) {
    const size_t a = St < Sh ? St : Sh;
    const size_t b = St < Sh ? Sh : St;
    const size_t cut[4] = {0, a, b, n};
    for (int k = 0; k < 3; ++k) {
        const size_t j0 = cut[k], j1 = cut[k + 1];
        if (j0 >= j1) continue;
        const float* tp = t + (j0 < St ? j0 + n - St : j0 - St);
        const float* hp = h + (j0 < Sh ? j0 + n - Sh : j0 - Sh);
        float* op = o + j0;
        const size_t m = j1 - j0;
        for (size_t j = 0; j < m; ++j) op[j] = tp[j] + hp[j];
    }
}

static inline void add_rot_f64(double* restrict o, const double* restrict t, size_t St,
                               const double* restrict h, size_t Sh, size_t n) {
    const size_t a = St < Sh ? St : Sh;
    const size_t b = St < Sh ? Sh : St;
    const size_t cut[4] = {0, a, b, n};
    for (int k = 0; k < 3; ++k) {
        const size_t j0 = cut[k], j1 = cut[k + 1];
        if (j0 >= j1) continue;
        const double* tp = t + (j0 < St ? j0 + n - St : j0 - St);
        const double* hp = h + (j0 < Sh ? j0 + n - Sh : j0 - Sh);
        double* op = o + j0;
        const size_t m = j1 - j0;
        for (size_t j = 0; j < m; ++j) op[j] = tp[j] + hp[j];
    }
}

/*
 * Freq-only time merge, kernels.cpp:118-146 (shift_add_linear_with_buffer)
 * and 722-827: out[s,c,q,j] = in[2s,idx,q,j] + in[2s+1,idx,q,(j-S) mod n].
 */
void orc_ffa_iter_freq(const float* in, float* out, const orc_coord_freq* coords,
                       size_t ncur, size_t nprev, size_t nseg, size_t nbins,
                       int nthreads) {
    const size_t row = 2 * nbins;
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for collapse(2) num_threads(nt) schedule(static)
    for (size_t s = 0; s < nseg; ++s) {
        for (size_t c = 0; c < ncur; ++c) {
            const float* t = in + (2 * s) * nprev * row + (size_t)coords[c].idx * row;
            const float* h = in + (2 * s + 1) * nprev * row + (size_t)coords[c].idx * row;
            float* o = out + s * ncur * row + c * row;
            const size_t S = shift_bins(coords[c].shift, nbins);
            for (size_t q = 0; q < 2; ++q)
                add_rot_f32(o + q * nbins, t + q * nbins, 0, h + q * nbins, S, nbins);
        }
    }
}

/*
 * Poly (accel/jerk) time merge, kernels.cpp:71-112
 * (shift_add_binary_with_buffer) and 829-940.
 */
void orc_ffa_iter_poly(const float* in, float* out, const orc_coord* coords,
                       size_t ncur, size_t nprev, size_t nseg, size_t nbins,
                       int nthreads) {
    const size_t row = 2 * nbins;
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for collapse(2) num_threads(nt) schedule(static)
    for (size_t s = 0; s < nseg; ++s) {
        for (size_t c = 0; c < ncur; ++c) {
            const float* t = in + (2 * s) * nprev * row + (size_t)coords[c].i_tail * row;
            const float* h = in + (2 * s + 1) * nprev * row + (size_t)coords[c].i_head * row;
            float* o = out + s * ncur * row + c * row;
            const size_t St = shift_bins(coords[c].shift_tail, nbins);
            const size_t Sh = shift_bins(coords[c].shift_head, nbins);
            for (size_t q = 0; q < 2; ++q)
                add_rot_f32(o + q * nbins, t + q * nbins, St, h + q * nbins, Sh, nbins);
        }
    }
}

/* FP64 twins of the two merges (FP64 oracle fold for precision studies). */
void orc_ffa_iter_freq_f64(const double* in, double* out, const orc_coord_freq* coords,
                           size_t ncur, size_t nprev, size_t nseg, size_t nbins,
                           int nthreads) {
    const size_t row = 2 * nbins;
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for collapse(2) num_threads(nt) schedule(static)
    for (size_t s = 0; s < nseg; ++s)
        for (size_t c = 0; c < ncur; ++c) {
            const double* t = in + (2 * s) * nprev * row + (size_t)coords[c].idx * row;
            const double* h = in + (2 * s + 1) * nprev * row + (size_t)coords[c].idx * row;
            double* o = out + s * ncur * row + c * row;
            const size_t S = shift_bins(coords[c].shift, nbins);
            for (size_t q = 0; q < 2; ++q)
                add_rot_f64(o + q * nbins, t + q * nbins, 0, h + q * nbins, S, nbins);
        }
}

void orc_ffa_iter_poly_f64(const double* in, double* out, const orc_coord* coords,
                           size_t ncur, size_t nprev, size_t nseg, size_t nbins,
                           int nthreads) {
    const size_t row = 2 * nbins;
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for collapse(2) num_threads(nt) schedule(static)
    for (size_t s = 0; s < nseg; ++s)
        for (size_t c = 0; c < ncur; ++c) {
            const double* t = in + (2 * s) * nprev * row + (size_t)coords[c].i_tail * row;
            const double* h = in + (2 * s + 1) * nprev * row + (size_t)coords[c].i_head * row;
            double* o = out + s * ncur * row + c * row;
            const size_t St = shift_bins(coords[c].shift_tail, nbins);
            const size_t Sh = shift_bins(coords[c].shift_head, nbins);
            for (size_t q = 0; q < 2; ++q)
                add_rot_f64(o + q * nbins, t + q * nbins, St, h + q * nbins, Sh, nbins);
        }
}

/*
 * Fourier brute fold in FP64, the textbook DFT that kernels.cpp:582-720 and
 * 1156-1242 approximate with an FP32 phasor recurrence:
 *   fold[s,f,q,m] = sum_k x_q[s*seglen+k] * exp(-2 pi i m f (k*tsamp - t_ref))
 * for m = 0..nbins/2.  fold is interleaved (re, im) float pairs
 * [nseg, nfreqs, 2, nbins_f, 2].  The phase f*t is reduced mod 1 in FP64
 * before the harmonic multiply.
 */
void orc_brute_fold_fourier(const float* ts_e, const float* ts_v, size_t nsamps,
                            const double* freqs, size_t nfreqs, size_t seglen,
                            size_t nbins, double tsamp, double t_ref, float* fold,
                            int nthreads) {
    const size_t nseg = nsamps / seglen;
    const size_t nbf = nbins / 2 + 1;
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for collapse(2) num_threads(nt) schedule(static)
    for (size_t s = 0; s < nseg; ++s) {
        for (size_t f = 0; f < nfreqs; ++f) {
            const float* e = ts_e + s * seglen;
            const float* v = ts_v + s * seglen;
            float* out = fold + ((s * nfreqs + f) * 2 * nbf) * 2;
            for (size_t m = 0; m < nbf; ++m) {
                double er = 0, ei = 0, vr = 0, vi = 0;
                for (size_t k = 0; k < seglen; ++k) {
                    const double ph = freqs[f] * (((double)k * tsamp) - t_ref);
                    double fr = ph - floor(ph);
                    double pm = fr * (double)m;
                    pm = pm - floor(pm);
                    const double c = cos(2.0 * ORC_PI * pm);
                    const double sn = -sin(2.0 * ORC_PI * pm);
                    er += (double)e[k] * c;
                    ei += (double)e[k] * sn;
                    vr += (double)v[k] * c;
                    vi += (double)v[k] * sn;
                }
                out[2 * m] = (float)er;
                out[2 * m + 1] = (float)ei;
                out[2 * (nbf + m)] = (float)vr;
                out[2 * (nbf + m) + 1] = (float)vi;
            }
        }
    }
}

/* exp(-2 pi i k s / n) in FP64 with the product k*s reduced mod n. */
static inline void twiddle(size_t k, float s, size_t n, double* c, double* sn) {
    double a = (double)k * (double)s;
    a = fmod(a, (double)n);
    const double ang = -2.0 * ORC_PI * a / (double)n;
    *c = cos(ang);
    *sn = sin(ang);
}

/*
 * Fourier merges in FP64 (kernels.cpp:241-453, 942-1128):
 *   out_k = tail_k * w(s_t)^k + head_k * w(s_h)^k,  w(s) = exp(-2 pi i s / n)
 * (freq-only: tail factor 1).  Input/output interleaved complex64
 * [nseg, ncoords, 2, nbins_f]; arithmetic in FP64, one rounding per output.
 */
void orc_ffa_complex_iter_poly(const float* in, float* out, const orc_coord* coords,
                               size_t ncur, size_t nprev, size_t nseg, size_t nbins,
                               int nthreads) {
    const size_t nbf = nbins / 2 + 1;
    const size_t row = 2 * nbf * 2; /* floats per coord row */
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for collapse(2) num_threads(nt) schedule(static)
    for (size_t s = 0; s < nseg; ++s)
        for (size_t c = 0; c < ncur; ++c) {
            const float* t = in + (2 * s) * nprev * row + (size_t)coords[c].i_tail * row;
            const float* h = in + (2 * s + 1) * nprev * row + (size_t)coords[c].i_head * row;
            float* o = out + s * ncur * row + c * row;
            for (size_t k = 0; k < nbf; ++k) {
                double ct, st, ch, sh;
                twiddle(k, coords[c].shift_tail, nbins, &ct, &st);
                twiddle(k, coords[c].shift_head, nbins, &ch, &sh);
                for (size_t q = 0; q < 2; ++q) {
                    const size_t i = 2 * (q * nbf + k);
                    const double tr = t[i], ti = t[i + 1], hr = h[i], hi = h[i + 1];
                    o[i] = (float)((tr * ct - ti * st) + (hr * ch - hi * sh));
                    o[i + 1] = (float)((tr * st + ti * ct) + (hr * sh + hi * ch));
                }
            }
        }
}

void orc_ffa_complex_iter_freq(const float* in, float* out, const orc_coord_freq* coords,
                               size_t ncur, size_t nprev, size_t nseg, size_t nbins,
                               int nthreads) {
    const size_t nbf = nbins / 2 + 1;
    const size_t row = 2 * nbf * 2;
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for collapse(2) num_threads(nt) schedule(static)
    for (size_t s = 0; s < nseg; ++s)
        for (size_t c = 0; c < ncur; ++c) {
            const float* t = in + (2 * s) * nprev * row + (size_t)coords[c].idx * row;
            const float* h = in + (2 * s + 1) * nprev * row + (size_t)coords[c].idx * row;
            float* o = out + s * ncur * row + c * row;
            for (size_t k = 0; k < nbf; ++k) {
                double ch, sh;
                twiddle(k, coords[c].shift, nbins, &ch, &sh);
                for (size_t q = 0; q < 2; ++q) {
                    const size_t i = 2 * (q * nbf + k);
                    const double hr = h[i], hi = h[i + 1];
                    o[i] = (float)((double)t[i] + (hr * ch - hi * sh));
                    o[i + 1] = (float)((double)t[i + 1] + (hr * sh + hi * ch));
                }
            }
        }
}

/*
 * Inverse real DFT with 1/n normalisation (the C2R transform of
 * fft_cuda.cu:357-448 / fft.cpp:503-620): for each of nfft rows of nbins_f
 * complex values, x_j = (1/n) [X_0 + 2 sum_{k=1}^{ceil(n/2)-1} Re(X_k e^{2 pi i jk/n})
 * + X_{n/2} (-1)^j (n even)]; imaginary parts of X_0 and X_{n/2} ignored.
 */
void orc_irfft(const float* in, float* out, size_t nfft, size_t nbins, int nthreads) {
    const size_t nbf = nbins / 2 + 1;
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t r = 0; r < nfft; ++r) {
        const float* X = in + r * nbf * 2;
        float* x = out + r * nbins;
        for (size_t j = 0; j < nbins; ++j) {
            double acc = X[0];
            const size_t kmax = (nbins % 2 == 0) ? nbins / 2 : nbins / 2 + 1;
            for (size_t k = 1; k < kmax; ++k) {
                const double a = 2.0 * ORC_PI * (double)((j * k) % nbins) / (double)nbins;
                acc += 2.0 * ((double)X[2 * k] * cos(a) - (double)X[2 * k + 1] * sin(a));
            }
            if (nbins % 2 == 0) acc += (double)X[2 * (nbins / 2)] * ((j % 2) ? -1.0 : 1.0);
            x[j] = (float)(acc / (double)nbins);
        }
    }
}

/* utils.cpp:33-67 circular_prefix_sum. */
void orc_circular_prefix_sum(const float* x, float* out, size_t nbins, size_t nsum) {
    if (nbins == 0 || nsum == 0) return;
    out[0] = x[0];
    for (size_t i = 1; i < nbins && i < nsum; ++i) out[i] = out[i - 1] + x[i];
    if (nsum <= nbins) return;
    const float last = out[nbins - 1];
    const size_t first_end = nsum < 2 * nbins ? nsum : 2 * nbins;
    for (size_t i = nbins; i < first_end; ++i) out[i] = out[i - nbins] + last;
    for (size_t i = 2 * nbins; i < nsum; ++i)
        out[i] = out[i % nbins] + ((float)(i / nbins) * last);
}

/* utils.cpp:22-31 diff_max. */
float orc_diff_max(const float* x, const float* y, size_t n) {
    float m = -3.402823466e+38f;
    for (size_t i = 0; i < n; ++i) {
        const float d = x[i] - y[i];
        m = d > m ? d : m;
    }
    return m;
}

/*
 * score.cpp:95-172 snr_boxcar_impl<Is3D=true>: folds [nprof, 2, nbins];
 * scores [nprof, nwidths] (find_max=0) or [nprof] (find_max=1).
 */
void orc_snr_boxcar_3d(const float* folds, size_t nprof, size_t nbins,
                       const uint64_t* widths, size_t nw, float* scores,
                       int find_max, int nthreads) {
    size_t wmax = 0;
    for (size_t i = 0; i < nw; ++i) wmax = widths[i] > wmax ? widths[i] : wmax;
    float* hv = (float*)malloc(nw * sizeof(float));
    float* bv = (float*)malloc(nw * sizeof(float));
    for (size_t i = 0; i < nw; ++i) {
        const size_t w = widths[i];
        hv[i] = sqrtf((float)(nbins - w) / (float)(nbins * w));
        bv[i] = (float)w * hv[i] / (float)(nbins - w);
    }
    const int nt = clamp_threads(nthreads);
#pragma omp parallel num_threads(nt)
    {
        float* work = (float*)malloc(nbins * sizeof(float));
        float* psum = (float*)malloc((nbins + wmax) * sizeof(float));
#pragma omp for schedule(static)
        for (size_t p = 0; p < nprof; ++p) {
            const float* e = folds + p * 2 * nbins;
            const float* v = e + nbins;
            for (size_t j = 0; j < nbins; ++j) {
                const float inv = 1.0f / sqrtf(v[j]);
                work[j] = e[j] * inv;
            }
            orc_circular_prefix_sum(work, psum, nbins, nbins + wmax);
            const float sum = psum[nbins - 1];
            float best = -3.402823466e+38f;
            for (size_t i = 0; i < nw; ++i) {
                const float dmax = orc_diff_max(psum + widths[i], psum, nbins);
                const float snr = ((hv[i] + bv[i]) * dmax) - (bv[i] * sum);
                if (find_max) {
                    best = snr > best ? snr : best;
                } else {
                    scores[p * nw + i] = snr;
                }
            }
            if (find_max) scores[p] = best;
        }
        free(work);
        free(psum);
    }
    free(hv);
    free(bv);
}

/* FP64 boxcar S/N on an FP64 fold: the reference value for precision studies. */
void orc_snr_boxcar_3d_f64(const double* folds, size_t nprof, size_t nbins,
                           const uint64_t* widths, size_t nw, double* scores,
                           int nthreads) {
    const int nt = clamp_threads(nthreads);
#pragma omp parallel num_threads(nt)
    {
        double* psum = (double*)malloc(2 * nbins * sizeof(double));
#pragma omp for schedule(static)
        for (size_t p = 0; p < nprof; ++p) {
            const double* e = folds + p * 2 * nbins;
            const double* v = e + nbins;
            double acc = 0.0;
            for (size_t j = 0; j < nbins; ++j) {
                acc += e[j] / sqrt(v[j]);
                psum[j] = acc;
            }
            for (size_t j = nbins; j < 2 * nbins; ++j) psum[j] = psum[j - nbins] + acc;
            for (size_t i = 0; i < nw; ++i) {
                const size_t w = widths[i];
                const double h = sqrt((double)(nbins - w) / (double)(nbins * w));
                const double b = (double)w * h / (double)(nbins - w);
                double dmax = -1e300;
                for (size_t j = 0; j < nbins; ++j) {
                    const double d = psum[j + w] - psum[j];
                    dmax = d > dmax ? d : dmax;
                }
                scores[p * nw + i] = (h + b) * dmax - b * acc;
            }
        }
        free(psum);
    }
}

/*
 * Reduced-precision fold storage (the precision addition; no reference
 * counterpart): x[i] = round(x[i] * scale) to IEEE binary16 (storage 1) or
 * bfloat16 (storage 2), round to nearest even, returned as the exactly equal
 * float.  scale is a power of two, so x * scale is exact.  binary16 overflow
 * (|y| >= 65520) gives inf; NaN stays NaN.
 */
static inline float orc_round_f16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    const uint32_t sign = u & 0x80000000u;
    const uint32_t a = u & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) return x;                 /* inf, NaN */
    if (a >= 0x477FF000u) return sign ? -INFINITY : INFINITY;
    if (a >= 0x38800000u) {                          /* normal binary16: keep 10 bits */
        uint32_t r = a + 0x0FFFu + ((a >> 13) & 1u);
        r &= ~0x1FFFu;
        r |= sign;
        float y;
        memcpy(&y, &r, 4);
        return y;
    }
    /* subnormal binary16: multiples of 2^-24; |x| * 2^24 < 2^10 is exact */
    const float q = rintf(fabsf(x) * 16777216.0f) * (1.0f / 16777216.0f);
    return sign ? -q : q;
}

static inline float orc_round_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return x;  /* NaN */
    const uint64_t w = (uint64_t)u + 0x7FFFu + ((u >> 16) & 1u);
    uint32_t r = (uint32_t)((w >> 16) << 16);
    float y;
    memcpy(&y, &r, 4);
    return y;
}

void orc_round_storage(float* x, size_t n, float scale, int storage, int nthreads) {
    const int nt = clamp_threads(nthreads);
    if (storage == 1) {
#pragma omp parallel for num_threads(nt) schedule(static)
        for (size_t i = 0; i < n; ++i) x[i] = orc_round_f16(x[i] * scale);
    } else if (storage == 2) {
#pragma omp parallel for num_threads(nt) schedule(static)
        for (size_t i = 0; i < n; ++i) x[i] = orc_round_bf16(x[i] * scale);
    } else {
#pragma omp parallel for num_threads(nt) schedule(static)
        for (size_t i = 0; i < n; ++i) x[i] = x[i] * scale;
    }
}

/* Block floating-point storage (bfp16): each component row of nbins FP32 values is stored as
 * nbins int16 mantissas m and one exponent E, value = m * 2^E.  E = max(ex(M) - 14, -126) with
 * M = max |x| over the row and ex(M) the unbiased exponent field of M (-127 for zero or
 * subnormal M), so M * 2^-E < 2^15; m = rint(x * 2^-E) (ties to even) clamped to +-32767.
 * x * 2^-E and m * 2^E are exact in FP32, so the single rounding is the rint.  In place: x
 * becomes the decoded value m * 2^E.  Finite inputs only. */
static inline float orc_pow2f(int e) {
    uint32_t b = (uint32_t)(e + 127) << 23;
    float f;
    memcpy(&f, &b, sizeof f);
    return f;
}

int orc_bfp16_exponent(float mx) {
    uint32_t b;
    memcpy(&b, &mx, sizeof b);
    const int ex = (int)((b >> 23) & 0xFFu) - 127;
    const int e = ex - 14;
    return e < -126 ? -126 : e;
}

void orc_round_bfp16(float* x, size_t nrows, size_t nbins, int nthreads) {
    const int nt = clamp_threads(nthreads);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t r = 0; r < nrows; ++r) {
        float* row = x + r * nbins;
        float mx = 0.0f;
        for (size_t j = 0; j < nbins; ++j) {
            const float a = fabsf(row[j]);
            mx = a > mx ? a : mx;
        }
        const int e = orc_bfp16_exponent(mx);
        const float inv = orc_pow2f(-e);
        const float s = orc_pow2f(e);
        for (size_t j = 0; j < nbins; ++j) {
            float m = rintf(row[j] * inv);
            m = m > 32767.0f ? 32767.0f : (m < -32767.0f ? -32767.0f : m);
            row[j] = m * s;
        }
    }
}
