// Seeds of the pruning stage from final-level FFA folds (kernels_cuda.cu:16-277,
// kernels.cpp:1328-1452).  Rows are [e: W][v: W] with W = nbins floats (time)
// or nbf = nbins/2+1 complex values (Fourier).
//
//   linear:  out[l] = tree[phys(l)] + R_{s_l}(ffa[i_l])          (valid leaves only)
//   ascend:  out[l] = sum_{g=0}^{G-1} R_{s_gl}(ffa[seg_g][i_gl])  (every leaf)
//
// R_s is the right rotation by the integer shift u32(s + 0.5f) (nbins -> 0) in
// the time domain, and the product with exp(-2 pi i k s / nbins) over the
// harmonics k for the fractional shift s in the Fourier domain.  phys(l) is the
// circular tree index i + start, less `capacity` once when it reaches it.
//
// One lane per (leaf, bin or harmonic) handles both components, so a wave
// reads and writes contiguous runs of a row; with nbins % 4 == 0 (and 16-byte
// aligned buffers) a time-domain lane owns 4 bins as float4 and keeps several
// leaves' (or segments') loads in flight.  The time-domain sums keep the
// reference's order (0 + a_0 + a_1 + ...), so they are bit-identical to its
// CPU kernels; the Fourier twiddle is the accurate FP64-reduced one of the
// merges.  The reference takes no row counts and trusts its indices; here an
// index beyond the rows given writes NaN to that output bin instead of
// reading out of bounds.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>

#include "kernels.hpp"

namespace loki_amd::kern {
namespace {

constexpr int kBlock = 256;

unsigned seed_grid(uint64_t total) {
    const uint64_t want = (total + kBlock - 1) / kBlock;
    const uint64_t cap  = 65536ULL * 8;
    return static_cast<unsigned>(want < cap ? (want == 0 ? 1 : want) : cap);
}

/// kernels_cuda.cu:41-44: u32(s + 0.5f) with nbins -> 0; a shift beyond one
/// turn (no valid phase gives one) is taken mod nbins rather than indexing
/// before the row.
__device__ __forceinline__ uint32_t seed_shift(float s, uint32_t nbins) {
    uint32_t k = static_cast<uint32_t>(s + 0.5F);
    if (k == nbins) {
        k = 0;
    }
    return k < nbins ? k : k % nbins;
}

/// exp(-2 pi i d), d in turns reduced in FP64 (as the Fourier merges).
__device__ __forceinline__ float2 seed_twiddle(double d) {
    d -= floor(d);
    if (d > 0.5) {
        d -= 1.0;
    }
    const float hi = static_cast<float>(d);
    const float lo = static_cast<float>(d - static_cast<double>(hi));
    float sn, cs;
    sincospif(2.0F * hi, &sn, &cs);
    const float dl = 6.28318530717958647692F * lo;
    return make_float2(fmaf(-dl, sn, cs), -fmaf(dl, cs, sn));
}

__device__ __forceinline__ float2 seed_cmul(float2 a, float2 w) {
    return make_float2(fmaf(a.x, w.x, -(a.y * w.y)), fmaf(a.x, w.y, a.y * w.x));
}

template <typename IDX>
__global__ void __launch_bounds__(kBlock)
k_seed_linear(const float* __restrict__ tree, const uint32_t* __restrict__ itree,
              const uint8_t* __restrict__ mask, const float* __restrict__ ffa,
              const uint32_t* __restrict__ iffa, const float* __restrict__ shift,
              float* __restrict__ out, uint32_t nbins, IDX nleaves, uint32_t start,
              uint32_t capacity, uint64_t ffa_rows) {
    const IDX total = nleaves * nbins;
    const float nan = std::numeric_limits<float>::quiet_NaN();
    for (IDX t = static_cast<IDX>(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += static_cast<IDX>(gridDim.x) * kBlock) {
        const IDX leaf   = t / nbins;
        const uint32_t j = static_cast<uint32_t>(t - leaf * nbins);
        if (mask != nullptr && mask[leaf] == 0) {
            continue;
        }
        float* o = out + static_cast<uint64_t>(leaf) * 2 * nbins;
        uint32_t tr = itree[leaf] + start;
        if (tr >= capacity) {
            tr -= capacity;
        }
        const uint32_t fr = iffa[leaf];
        if (tr >= capacity || fr >= ffa_rows) {
            o[j]         = nan;
            o[nbins + j] = nan;
            continue;
        }
        const uint32_t S   = seed_shift(shift[leaf], nbins);
        const uint32_t src = j >= S ? j - S : j + nbins - S;
        const float* tp = tree + static_cast<uint64_t>(tr) * 2 * nbins;
        const float* fp = ffa + static_cast<uint64_t>(fr) * 2 * nbins;
        o[j]         = tp[j] + fp[src];
        o[nbins + j] = tp[nbins + j] + fp[nbins + src];
    }
}

template <typename IDX>
__global__ void __launch_bounds__(kBlock)
k_seed_ascend(const float* __restrict__ ffa, const uint32_t* __restrict__ iseg,
              const uint32_t* __restrict__ iffa, const float* __restrict__ shift,
              float* __restrict__ out, uint32_t nbins, uint64_t ncoords_init, IDX nleaves,
              uint32_t nsegments, uint64_t ffa_segments) {
    const IDX total = nleaves * nbins;
    for (IDX t = static_cast<IDX>(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += static_cast<IDX>(gridDim.x) * kBlock) {
        const IDX leaf   = t / nbins;
        const uint32_t j = static_cast<uint32_t>(t - leaf * nbins);
        float ae = 0.0F, av = 0.0F;
        for (uint32_t g = 0; g < nsegments; ++g) {
            const uint64_t ls = static_cast<uint64_t>(g) * nleaves + leaf;
            const uint32_t fi = iffa[ls];
            const uint32_t si = iseg[g];
            if (fi >= ncoords_init || si >= ffa_segments) {
                ae = av = std::numeric_limits<float>::quiet_NaN();
                break;
            }
            const uint32_t S   = seed_shift(shift[ls], nbins);
            const uint32_t src = j >= S ? j - S : j + nbins - S;
            const float* fp = ffa + (static_cast<uint64_t>(si) * ncoords_init + fi) * 2 * nbins;
            ae += fp[src];
            av += fp[nbins + src];
        }
        float* o     = out + static_cast<uint64_t>(leaf) * 2 * nbins;
        o[j]         = ae;
        o[nbins + j] = av;
    }
}

/// Elements r..r+3 of one component row rotated right by S (out[j] =
/// row[(j - S) mod n]) for j = 4*q..4*q+3, n a multiple of 4: the four
/// sources lie in at most two aligned float4 chunks.
__device__ __forceinline__ float4 seed_rot4(const float4* __restrict__ row, uint32_t q,
                                            uint32_t S, uint32_t n, uint32_t c4) {
    uint32_t e0 = 4 * q + n - S;
    e0          = e0 >= n ? e0 - n : e0;
    const uint32_t c0 = e0 >> 2;
    const uint32_t c1 = c0 + 1 == c4 ? 0 : c0 + 1;
    const float4 a    = row[c0];
    const float4 b    = row[c1];
    const uint32_t r  = e0 & 3U;
    float4 o;
    o.x = r == 0 ? a.x : (r == 1 ? a.y : (r == 2 ? a.z : a.w));
    o.y = r == 0 ? a.y : (r == 1 ? a.z : (r == 2 ? a.w : b.x));
    o.z = r == 0 ? a.z : (r == 1 ? a.w : (r == 2 ? b.x : b.y));
    o.w = r == 0 ? a.w : (r == 1 ? b.x : (r == 2 ? b.y : b.z));
    return o;
}

__device__ __forceinline__ float4 add4(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

/// nbins % 4 == 0: lane per (leaf, 4 bins) of both components; each lane
/// takes LPL leaves per pass with all their loads issued before the adds.
template <int LPL>
__global__ void __launch_bounds__(kBlock)
k_seed_linear4(const float* __restrict__ tree, const uint32_t* __restrict__ itree,
               const uint8_t* __restrict__ mask, const float* __restrict__ ffa,
               const uint32_t* __restrict__ iffa, const float* __restrict__ shift,
               float* __restrict__ out, uint32_t nbins, uint32_t nleaves, uint32_t start,
               uint32_t capacity, uint64_t ffa_rows) {
    const uint32_t c4    = nbins / 4;
    const uint64_t total = static_cast<uint64_t>(nleaves) * c4;
    const uint64_t step  = static_cast<uint64_t>(gridDim.x) * kBlock;
    const float nan      = std::numeric_limits<float>::quiet_NaN();
    for (uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; t0 < total;
         t0 += step * LPL) {
        float4 te[LPL], tv[LPL], he[LPL], hv[LPL];
        uint32_t leaf[LPL], q[LPL];
        bool live[LPL], bad[LPL];
#pragma unroll
        for (int u = 0; u < LPL; ++u) {
            const uint64_t t = t0 + u * step;
            live[u] = t < total;
            bad[u]  = false;
            leaf[u] = live[u] ? static_cast<uint32_t>(t / c4) : 0;
            q[u]    = live[u] ? static_cast<uint32_t>(t - static_cast<uint64_t>(leaf[u]) * c4) : 0;
            if (live[u] && mask != nullptr && mask[leaf[u]] == 0) {
                live[u] = false;
            }
            if (live[u]) {
                uint32_t tr = itree[leaf[u]] + start;
                if (tr >= capacity) {
                    tr -= capacity;
                }
                const uint32_t fr = iffa[leaf[u]];
                bad[u] = tr >= capacity || fr >= ffa_rows;
                if (!bad[u]) {
                    const uint32_t S = seed_shift(shift[leaf[u]], nbins);
                    const float4* tp =
                        reinterpret_cast<const float4*>(tree + static_cast<uint64_t>(tr) * 2 * nbins);
                    const float4* fp =
                        reinterpret_cast<const float4*>(ffa + static_cast<uint64_t>(fr) * 2 * nbins);
                    te[u] = tp[q[u]];
                    tv[u] = tp[c4 + q[u]];
                    he[u] = seed_rot4(fp, q[u], S, nbins, c4);
                    hv[u] = seed_rot4(fp + c4, q[u], S, nbins, c4);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < LPL; ++u) {
            if (!live[u]) {
                continue;
            }
            float4* o = reinterpret_cast<float4*>(out + static_cast<uint64_t>(leaf[u]) * 2 * nbins);
            if (bad[u]) {
                o[q[u]]      = make_float4(nan, nan, nan, nan);
                o[c4 + q[u]] = make_float4(nan, nan, nan, nan);
            } else {
                o[q[u]]      = add4(te[u], he[u]);
                o[c4 + q[u]] = add4(tv[u], hv[u]);
            }
        }
    }
}

/// nbins % 4 == 0: lane per (leaf, 4 bins); the segments are taken SPL at a
/// time with their index, shift and row loads issued together, and added in
/// segment order.
template <int SPL>
__global__ void __launch_bounds__(kBlock)
k_seed_ascend4(const float* __restrict__ ffa, const uint32_t* __restrict__ iseg,
               const uint32_t* __restrict__ iffa, const float* __restrict__ shift,
               float* __restrict__ out, uint32_t nbins, uint64_t ncoords_init, uint32_t nleaves,
               uint32_t nsegments, uint64_t ffa_segments) {
    const uint32_t c4    = nbins / 4;
    const uint64_t total = static_cast<uint64_t>(nleaves) * c4;
    for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint32_t leaf = static_cast<uint32_t>(t / c4);
        const uint32_t q    = static_cast<uint32_t>(t - static_cast<uint64_t>(leaf) * c4);
        float4 ae = make_float4(0.0F, 0.0F, 0.0F, 0.0F), av = ae;
        bool bad = false;
        for (uint32_t g0 = 0; g0 < nsegments && !bad; g0 += SPL) {
            float4 he[SPL], hv[SPL];
            bool use[SPL];
#pragma unroll
            for (int u = 0; u < SPL; ++u) {
                const uint32_t g = g0 + u;
                use[u]           = g < nsegments;
                if (use[u]) {
                    const uint64_t ls = static_cast<uint64_t>(g) * nleaves + leaf;
                    const uint32_t fi = iffa[ls];
                    const uint32_t si = iseg[g];
                    if (fi >= ncoords_init || si >= ffa_segments) {
                        bad    = true;
                        use[u] = false;
                    } else {
                        const uint32_t S = seed_shift(shift[ls], nbins);
                        const float4* fp = reinterpret_cast<const float4*>(
                            ffa + (static_cast<uint64_t>(si) * ncoords_init + fi) * 2 * nbins);
                        he[u] = seed_rot4(fp, q, S, nbins, c4);
                        hv[u] = seed_rot4(fp + c4, q, S, nbins, c4);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < SPL; ++u) {
                if (use[u]) {
                    ae = add4(ae, he[u]);
                    av = add4(av, hv[u]);
                }
            }
        }
        float4* o = reinterpret_cast<float4*>(out + static_cast<uint64_t>(leaf) * 2 * nbins);
        if (bad) {
            const float nan = std::numeric_limits<float>::quiet_NaN();
            ae = av = make_float4(nan, nan, nan, nan);
        }
        o[q]      = ae;
        o[c4 + q] = av;
    }
}

template <typename IDX>
__global__ void __launch_bounds__(kBlock)
k_seed_linear_c(const float2* __restrict__ tree, const uint32_t* __restrict__ itree,
                const uint8_t* __restrict__ mask, const float2* __restrict__ ffa,
                const uint32_t* __restrict__ iffa, const float* __restrict__ shift,
                float2* __restrict__ out, uint32_t nbf, uint32_t nbins, IDX nleaves,
                uint32_t start, uint32_t capacity, uint64_t ffa_rows) {
    const IDX total    = nleaves * nbf;
    const double inv_n = 1.0 / static_cast<double>(nbins);
    const float nan    = std::numeric_limits<float>::quiet_NaN();
    for (IDX t = static_cast<IDX>(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += static_cast<IDX>(gridDim.x) * kBlock) {
        const IDX leaf   = t / nbf;
        const uint32_t k = static_cast<uint32_t>(t - leaf * nbf);
        if (mask != nullptr && mask[leaf] == 0) {
            continue;
        }
        float2* o = out + static_cast<uint64_t>(leaf) * 2 * nbf;
        uint32_t tr = itree[leaf] + start;
        if (tr >= capacity) {
            tr -= capacity;
        }
        const uint32_t fr = iffa[leaf];
        if (tr >= capacity || fr >= ffa_rows) {
            o[k]       = make_float2(nan, nan);
            o[nbf + k] = make_float2(nan, nan);
            continue;
        }
        const float2 w = seed_twiddle(static_cast<double>(k) *
                                      static_cast<double>(shift[leaf]) * inv_n);
        const float2* tp = tree + static_cast<uint64_t>(tr) * 2 * nbf;
        const float2* fp = ffa + static_cast<uint64_t>(fr) * 2 * nbf;
        const float2 te = tp[k], tv = tp[nbf + k];
        const float2 he = seed_cmul(fp[k], w), hv = seed_cmul(fp[nbf + k], w);
        o[k]       = make_float2(te.x + he.x, te.y + he.y);
        o[nbf + k] = make_float2(tv.x + hv.x, tv.y + hv.y);
    }
}

template <typename IDX>
__global__ void __launch_bounds__(kBlock)
k_seed_ascend_c(const float2* __restrict__ ffa, const uint32_t* __restrict__ iseg,
                const uint32_t* __restrict__ iffa, const float* __restrict__ shift,
                float2* __restrict__ out, uint32_t nbf, uint32_t nbins, uint64_t ncoords_init,
                IDX nleaves, uint32_t nsegments, uint64_t ffa_segments) {
    const IDX total    = nleaves * nbf;
    const double inv_n = 1.0 / static_cast<double>(nbins);
    for (IDX t = static_cast<IDX>(blockIdx.x) * kBlock + threadIdx.x; t < total;
         t += static_cast<IDX>(gridDim.x) * kBlock) {
        const IDX leaf   = t / nbf;
        const uint32_t k = static_cast<uint32_t>(t - leaf * nbf);
        float2 ae = make_float2(0.0F, 0.0F), av = make_float2(0.0F, 0.0F);
        for (uint32_t g = 0; g < nsegments; ++g) {
            const uint64_t ls = static_cast<uint64_t>(g) * nleaves + leaf;
            const uint32_t fi = iffa[ls];
            const uint32_t si = iseg[g];
            if (fi >= ncoords_init || si >= ffa_segments) {
                const float nan = std::numeric_limits<float>::quiet_NaN();
                ae = av = make_float2(nan, nan);
                break;
            }
            const float2 w = seed_twiddle(static_cast<double>(k) *
                                          static_cast<double>(shift[ls]) * inv_n);
            const float2* fp =
                ffa + (static_cast<uint64_t>(si) * ncoords_init + fi) * 2 * nbf;
            const float2 he = seed_cmul(fp[k], w), hv = seed_cmul(fp[nbf + k], w);
            ae = make_float2(ae.x + he.x, ae.y + he.y);
            av = make_float2(av.x + hv.x, av.y + hv.y);
        }
        float2* o  = out + static_cast<uint64_t>(leaf) * 2 * nbf;
        o[k]       = ae;
        o[nbf + k] = av;
    }
}

/// float4 row access needs 16-byte aligned bases (rows of nbins % 4 == 0
/// floats keep that alignment).
bool aligned16(const void* a, const void* b, const void* c) {
    return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b) |
             reinterpret_cast<uintptr_t>(c)) & 15U) == 0;
}

/// 32-bit lane indices when the whole grid-stride range fits.
bool narrow(uint64_t total) {
    return total + static_cast<uint64_t>(seed_grid(total)) * kBlock < (1ULL << 32);
}

}  // namespace

void seed_linear(const float* tree, const uint32_t* itree, const uint8_t* mask, const float* ffa,
                 const uint32_t* iffa, const float* shift, float* out, uint32_t nbins,
                 uint64_t nleaves, uint32_t start, uint32_t capacity, uint64_t ffa_rows,
                 hipStream_t s) {
    const uint64_t total = nleaves * nbins;
    if (total == 0) {
        return;
    }
    if (nbins % 4 == 0 && nleaves < (1ULL << 32) && aligned16(tree, ffa, out)) {
        k_seed_linear4<4><<<seed_grid((total / 4 + 3) / 4), kBlock, 0, s>>>(
            tree, itree, mask, ffa, iffa, shift, out, nbins, static_cast<uint32_t>(nleaves), start,
            capacity, ffa_rows);
        return;
    }
    const unsigned g = seed_grid(total);
    if (narrow(total)) {
        k_seed_linear<uint32_t><<<g, kBlock, 0, s>>>(tree, itree, mask, ffa, iffa, shift, out,
                                                     nbins, static_cast<uint32_t>(nleaves),
                                                     start, capacity, ffa_rows);
    } else {
        k_seed_linear<uint64_t><<<g, kBlock, 0, s>>>(tree, itree, mask, ffa, iffa, shift, out,
                                                     nbins, nleaves, start, capacity, ffa_rows);
    }
}

void seed_ascend(const float* ffa, const uint32_t* iseg, const uint32_t* iffa,
                 const float* shift, float* out, uint32_t nbins, uint64_t ncoords_init,
                 uint64_t nleaves, uint32_t nsegments, uint64_t ffa_segments, hipStream_t s) {
    const uint64_t total = nleaves * nbins;
    if (total == 0) {
        return;
    }
    if (nbins % 4 == 0 && nleaves < (1ULL << 32) && aligned16(ffa, ffa, out)) {
        k_seed_ascend4<4><<<seed_grid(total / 4), kBlock, 0, s>>>(
            ffa, iseg, iffa, shift, out, nbins, ncoords_init, static_cast<uint32_t>(nleaves),
            nsegments, ffa_segments);
        return;
    }
    const unsigned g = seed_grid(total);
    if (narrow(total)) {
        k_seed_ascend<uint32_t><<<g, kBlock, 0, s>>>(ffa, iseg, iffa, shift, out, nbins,
                                                     ncoords_init,
                                                     static_cast<uint32_t>(nleaves), nsegments,
                                                     ffa_segments);
    } else {
        k_seed_ascend<uint64_t><<<g, kBlock, 0, s>>>(ffa, iseg, iffa, shift, out, nbins,
                                                     ncoords_init, nleaves, nsegments,
                                                     ffa_segments);
    }
}

void seed_linear_complex(const float2* tree, const uint32_t* itree, const uint8_t* mask,
                         const float2* ffa, const uint32_t* iffa, const float* shift,
                         float2* out, uint32_t nbf, uint32_t nbins, uint64_t nleaves,
                         uint32_t start, uint32_t capacity, uint64_t ffa_rows, hipStream_t s) {
    const uint64_t total = nleaves * nbf;
    if (total == 0) {
        return;
    }
    const unsigned g = seed_grid(total);
    if (narrow(total)) {
        k_seed_linear_c<uint32_t><<<g, kBlock, 0, s>>>(tree, itree, mask, ffa, iffa, shift, out,
                                                       nbf, nbins,
                                                       static_cast<uint32_t>(nleaves), start,
                                                       capacity, ffa_rows);
    } else {
        k_seed_linear_c<uint64_t><<<g, kBlock, 0, s>>>(tree, itree, mask, ffa, iffa, shift, out,
                                                       nbf, nbins, nleaves, start, capacity,
                                                       ffa_rows);
    }
}

void seed_ascend_complex(const float2* ffa, const uint32_t* iseg, const uint32_t* iffa,
                         const float* shift, float2* out, uint32_t nbf, uint32_t nbins,
                         uint64_t ncoords_init, uint64_t nleaves, uint32_t nsegments,
                         uint64_t ffa_segments, hipStream_t s) {
    const uint64_t total = nleaves * nbf;
    if (total == 0) {
        return;
    }
    const unsigned g = seed_grid(total);
    if (narrow(total)) {
        k_seed_ascend_c<uint32_t><<<g, kBlock, 0, s>>>(ffa, iseg, iffa, shift, out, nbf, nbins,
                                                       ncoords_init,
                                                       static_cast<uint32_t>(nleaves), nsegments,
                                                       ffa_segments);
    } else {
        k_seed_ascend_c<uint64_t><<<g, kBlock, 0, s>>>(ffa, iseg, iffa, shift, out, nbf, nbins,
                                                       ncoords_init, nleaves, nsegments,
                                                       ffa_segments);
    }
}

}  // namespace loki_amd::kern
