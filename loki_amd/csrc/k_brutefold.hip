// Time-domain brute fold without atomics.
//
// The reference CPU path (fold.cpp:113-142, kernels.cpp:550-580) sums each
// (freq, phase-bin) bucket in ascending sample order; its CUDA path uses
// float atomics and is order-nondeterministic (kernels_cuda.cu:662-773).
// Here the bucket lists are built once per plan (FP64 phases, one block per
// frequency) in a transposed layout list[f][k][b] = k-th sample of bin b, so
// that lanes holding consecutive bins read consecutive words at every step.
// The fold runs one lane per (freq, bin) and one to eight segments, summing
// each bucket in ascending order from samples staged in LDS -- bit-identical
// to the CPU order and deterministic.  Reduced-precision storage rounds each FP32 bucket
// sum once (storage.hpp).
#include <algorithm>
#include <array>
#include <mutex>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "grid.hpp"
#include "kernels.hpp"

namespace loki_amd::kern {
namespace {

constexpr int kChunk    = 4096;  // samples whose bins are counted in LDS at once
constexpr int kBlock    = 256;
constexpr int kPairsBlk = 2048;  // (freq, bin) pairs per fold block
constexpr int kLdsSamps = 4096;  // segments up to this length are staged in LDS

// Pass 1 (count): one block per frequency, thread b counts the samples of
// phase bin b; kmax[f] = largest bucket of frequency f.
__global__ void __launch_bounds__(1024)
k_bf_count(const double* __restrict__ freqs, uint32_t seglen, uint32_t nbins, double tsamp,
           double t_ref, uint32_t* __restrict__ counts, uint32_t* __restrict__ kmax) {
    __shared__ uint16_t s_bins[kChunk];
    __shared__ uint32_t s_max;
    const uint32_t f    = blockIdx.x;
    const double freq   = freqs[f];
    const uint32_t tid  = threadIdx.x;
    const uint32_t nthr = blockDim.x;
    if (tid == 0) {
        s_max = 0;
    }
    uint32_t count = 0;
    for (uint32_t base = 0; base < seglen; base += kChunk) {
        const uint32_t len = min(static_cast<uint32_t>(kChunk), seglen - base);
        __syncthreads();
        for (uint32_t k = tid; k < len; k += nthr) {
            const double t = (static_cast<double>(base + k) * tsamp) - t_ref;
            s_bins[k]      = static_cast<uint16_t>(grid::phase_bin(t, freq, nbins));
        }
        __syncthreads();
        if (tid < nbins) {
            for (uint32_t k = 0; k < len; ++k) {
                count += (s_bins[k] == tid) ? 1U : 0U;
            }
        }
    }
    if (tid < nbins) {
        counts[static_cast<uint64_t>(f) * nbins + tid] = count;
        atomicMax(&s_max, count);
    }
    __syncthreads();
    if (tid == 0) {
        kmax[f] = s_max;
    }
}

// Bucket offsets on the device (the allocation-free path): base[f] = sum over
// f' < f of kmax[f'] * nbins -- the host loop of build_brute_fold_lists --
// and *total the sum over all frequencies.  One block: each thread sums a
// contiguous range of frequencies, then an inclusive scan of the partials.
__global__ void __launch_bounds__(1024)
k_bf_offsets(const uint32_t* __restrict__ kmax, uint32_t nfreqs, uint32_t nbins,
             uint64_t* __restrict__ base, uint64_t* __restrict__ total) {
    __shared__ uint64_t s_part[1024];
    const uint32_t t   = threadIdx.x;
    const uint64_t per = (static_cast<uint64_t>(nfreqs) + 1023) / 1024;
    const uint64_t f0  = min(static_cast<uint64_t>(nfreqs), t * per);
    const uint64_t f1  = min(static_cast<uint64_t>(nfreqs), f0 + per);
    uint64_t sum = 0;
    for (uint64_t f = f0; f < f1; ++f) {
        sum += static_cast<uint64_t>(kmax[f]) * nbins;
    }
    s_part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint64_t v = t >= off ? s_part[t - off] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    uint64_t run = s_part[t] - sum;
    for (uint64_t f = f0; f < f1; ++f) {
        base[f] = run;
        run += static_cast<uint64_t>(kmax[f]) * nbins;
    }
    if (t == 1023) {
        *total = s_part[1023];
    }
}

// Pass 2 (fill): list[base[f] + k * nbins + b] = k-th sample (ascending) of
// bin b of frequency f.
__global__ void __launch_bounds__(1024)
k_bf_fill(const double* __restrict__ freqs, uint32_t seglen, uint32_t nbins, double tsamp,
          double t_ref, const uint64_t* __restrict__ base_f, uint32_t* __restrict__ list) {
    __shared__ uint16_t s_bins[kChunk];
    const uint32_t f    = blockIdx.x;
    const double freq   = freqs[f];
    const uint32_t tid  = threadIdx.x;
    const uint32_t nthr = blockDim.x;
    uint32_t* __restrict__ lst = list + base_f[f] + tid;
    uint64_t k_out             = 0;
    for (uint32_t base = 0; base < seglen; base += kChunk) {
        const uint32_t len = min(static_cast<uint32_t>(kChunk), seglen - base);
        __syncthreads();
        for (uint32_t k = tid; k < len; k += nthr) {
            const double t = (static_cast<double>(base + k) * tsamp) - t_ref;
            s_bins[k]      = static_cast<uint16_t>(grid::phase_bin(t, freq, nbins));
        }
        __syncthreads();
        if (tid < nbins) {
            for (uint32_t k = 0; k < len; ++k) {
                if (s_bins[k] == tid) {
                    lst[k_out * nbins] = base + k;
                    ++k_out;
                }
            }
        }
    }
}

// One lane per (segment, freq, bin) pair; consecutive lanes are consecutive
// bins, so list reads and output writes are contiguous.  USE_LDS: the
// segment's samples are staged in LDS (seglen <= kLdsSamps; longer segments
// take the run path above).
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
                                              TO* __restrict__ fold, float scale) {
    __shared__ float s_e[USE_LDS ? kLdsSamps : 1];
    __shared__ float s_v[USE_LDS ? kLdsSamps : 1];
    const uint64_t seg   = bid / blocks_per_seg;
    const uint32_t pbase = (bid % blocks_per_seg) * kPairsBlk;
    const uint32_t npair = nfreqs * nbins;
    const float* __restrict__ ge = ts_e + seg * seglen;
    const float* __restrict__ gv = ts_v + seg * seglen;
    if constexpr (USE_LDS) {
        for (uint32_t k = threadIdx.x; k < seglen; k += kBlock) {
            s_e[k] = ge[k];
            s_v[k] = gv[k];
        }
        __syncthreads();
    }
    const float* __restrict__ e = USE_LDS ? s_e : ge;
    const float* __restrict__ v = USE_LDS ? s_v : gv;
    const uint32_t pend = min(pbase + static_cast<uint32_t>(kPairsBlk), npair);
    for (uint32_t p = pbase + threadIdx.x; p < pend; p += kBlock) {
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
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < nf * nbins; i += kBlock) {
        acc[i] = make_float2(0.0F, 0.0F);
    }
    if (tid < nf) {
        s_geom[tid] = run_geom(freqs[f0 + tid], seglen, tsamp, t_ref, nbins);
    }
    for (uint32_t c0 = 0; c0 < seglen; c0 += kUnitChunk) {
        const uint32_t c1 = min(c0 + static_cast<uint32_t>(kUnitChunk), seglen);
        __syncthreads();  // previous chunk's units are done with s_ev and s_pref
        for (uint32_t k = c0 + tid; k < c1; k += kBlock) {
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
        for (uint32_t t = tid; t < total; t += kBlock) {
            uint32_t lo = 0, hi = nf;  // largest fi with s_pref[fi] <= t
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (s_pref[mid] <= t) {
                    lo = mid;
                } else {
                    hi = mid;
                }
            }
            const uint32_t fi  = lo;
            const int64_t u0   = s_ulo[fi] + static_cast<int64_t>(t - s_pref[fi]);
            const int64_t uhi  = s_uhi[fi];
            const RunGeom g    = s_geom[fi];
            const uint32_t b   = static_cast<uint32_t>(((u0 % nb) + nb) % nb);
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
                // 16-byte reads of two (e, v) pairs: fewer LDS instructions, and 8 lanes per
                // LDS pass instead of 16 (fewer bank collisions between the lanes' runs)
                const auto* s4 = reinterpret_cast<const float4*>(s_ev);
                for (; k + 8 <= z; k += 8) {
                    float4 y[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        y[i] = s4[(k - c0) / 2 + i];
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        ae = ae + y[i].x;
                        av = av + y[i].y;
                        ae = ae + y[i].z;
                        av = av + y[i].w;
                    }
                }
                for (; k < z; ++k) {
                    const float2 x = s_ev[k - c0];
                    ae = ae + x.x;
                    av = av + x.y;
                }
            }
            *slot = make_float2(ae, av);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < nf * nbins; i += kBlock) {
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

template <typename TO>
__global__ void __launch_bounds__(kBlock)
k_bf_units(const float* __restrict__ ts_e, const float* __restrict__ ts_v, uint32_t seglen,
           uint32_t nfreqs, uint32_t nbins, uint32_t blocks_per_seg,
           const double* __restrict__ freqs, double tsamp, double t_ref, TO* __restrict__ fold,
           float scale) {
    bf_units_block<TO>(blockIdx.x, ts_e, ts_v, seglen, nfreqs, nbins, blocks_per_seg, freqs,
                       tsamp, t_ref, fold, scale);
}

template <bool USE_LDS, typename TO>
__global__ void __launch_bounds__(kBlock)
k_bf_fold(const float* __restrict__ ts_e, const float* __restrict__ ts_v, uint32_t seglen,
          uint32_t nfreqs, uint32_t nbins, uint32_t blocks_per_seg,
          const uint32_t* __restrict__ list, const uint64_t* __restrict__ base_f,
          const uint32_t* __restrict__ counts, TO* __restrict__ fold, float scale) {
    bf_fold_block<USE_LDS, TO>(blockIdx.x, ts_e, ts_v, seglen, nfreqs, nbins, blocks_per_seg,
                               list, base_f, counts, fold, scale);
}

// Several segments per block: the bucket lists do not depend on the segment,
// so each list entry read from L2 serves NS gathers from NS staged segments
// (dynamic LDS: NS x seglen (e, v) pairs, one 8-byte LDS read per sample).
// Every bucket is still summed in ascending sample order, so the result is
// the one of k_bf_fold.
template <int NS, typename TO>
__global__ void __launch_bounds__(kBlock)
k_bf_fold_ms(const float* __restrict__ ts_e, const float* __restrict__ ts_v, uint32_t seglen,
             uint64_t nseg, uint32_t nfreqs, uint32_t nbins, uint32_t blocks_per_seg,
             const uint32_t* __restrict__ list, const uint64_t* __restrict__ base_f,
             const uint32_t* __restrict__ counts, TO* __restrict__ fold, float scale) {
    extern __shared__ float2 s_ev2[];  // [NS][seglen] (e, v)
    const uint64_t sg0   = static_cast<uint64_t>(blockIdx.x / blocks_per_seg) * NS;
    const uint32_t pbase = (blockIdx.x % blocks_per_seg) * kPairsBlk;
    const uint32_t nsv   = static_cast<uint32_t>(nseg - sg0 < NS ? nseg - sg0 : NS);
    for (uint32_t q = 0; q < nsv; ++q) {
        const float* __restrict__ ge = ts_e + (sg0 + q) * seglen;
        const float* __restrict__ gv = ts_v + (sg0 + q) * seglen;
        float2* __restrict__ l2      = s_ev2 + q * seglen;
        for (uint32_t k = threadIdx.x; k < seglen; k += kBlock) {
            l2[k] = make_float2(ge[k], gv[k]);
        }
    }
    __syncthreads();
    const uint32_t npair = nfreqs * nbins;
    const uint32_t pend  = min(pbase + static_cast<uint32_t>(kPairsBlk), npair);
    for (uint32_t p = pbase + threadIdx.x; p < pend; p += kBlock) {
        const uint32_t f   = p / nbins;
        const uint32_t b   = p - f * nbins;
        const uint32_t cnt = counts[p];
        const uint32_t* __restrict__ lst = list + base_f[f] + b;
        float se[NS], sv[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            se[q] = 0.0F;
            sv[q] = 0.0F;
        }
        uint32_t k = 0;
        for (; k + 2 <= cnt; k += 2) {
            const uint32_t i0 = lst[static_cast<uint64_t>(k) * nbins];
            const uint32_t i1 = lst[static_cast<uint64_t>(k + 1) * nbins];
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                const float2* __restrict__ l2 = s_ev2 + q * seglen;
                const float2 x0 = l2[i0];
                const float2 x1 = l2[i1];
                se[q] = se[q] + x0.x;
                sv[q] = sv[q] + x0.y;
                se[q] = se[q] + x1.x;
                sv[q] = sv[q] + x1.y;
            }
        }
        if (k < cnt) {
            const uint32_t i = lst[static_cast<uint64_t>(k) * nbins];
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                const float2 x = s_ev2[q * seglen + i];
                se[q] = se[q] + x.x;
                sv[q] = sv[q] + x.y;
            }
        }
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            if (static_cast<uint32_t>(q) < nsv) {
                TO* __restrict__ out = fold + ((sg0 + q) * nfreqs + f) * 2 * nbins;
                if constexpr (std::is_same_v<TO, float>) {
                    out[b]         = se[q];
                    out[nbins + b] = sv[q];
                } else {
                    out[b]         = narrow1<TO>(se[q] * scale);
                    out[nbins + b] = narrow1<TO>(sv[q] * scale);
                }
            }
        }
    }
}

// Grouped brute folds of many series (one task each, sorted by block0).
template <bool USE_LDS, typename TO>
__global__ void __launch_bounds__(kBlock)
k_bf_fold_grouped(const FoldTask* __restrict__ tasks, uint32_t ntasks) {
    uint32_t lo = 0;
    uint32_t hi = ntasks;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (tasks[mid].block0 <= blockIdx.x) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    const FoldTask& t = tasks[lo];
    bf_fold_block<USE_LDS, TO>(blockIdx.x - t.block0, t.ts_e, t.ts_v, t.seglen, t.nfreqs,
                               t.nbins, t.bps, t.list, t.base, t.counts,
                               static_cast<TO*>(t.fold), t.scale);
}

template <typename TO>
__global__ void __launch_bounds__(kBlock)
k_bf_units_grouped(const FoldTask* __restrict__ tasks, uint32_t ntasks) {
    uint32_t lo = 0;
    uint32_t hi = ntasks;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (tasks[mid].block0 <= blockIdx.x) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    const FoldTask& t = tasks[lo];
    bf_units_block<TO>(blockIdx.x - t.block0, t.ts_e, t.ts_v, t.seglen, t.nfreqs, t.nbins, t.bps,
                       t.freqs, t.tsamp, t.t_ref, static_cast<TO*>(t.fold), t.scale);
}

}  // namespace

BruteFoldLists build_brute_fold_lists(const double* freqs_dev, uint32_t nfreqs, uint32_t seglen,
                                      uint32_t nbins, double tsamp, double t_ref,
                                      hipStream_t s) {
    LOKI_REQUIRE(nbins >= 1 && nbins <= 1024, "brute fold supports 1 <= nbins <= 1024");
    LOKI_REQUIRE(nfreqs >= 1 && seglen >= 1, "brute fold: empty frequency or segment");
    LOKI_REQUIRE(tsamp > 0.0, "brute fold: tsamp must be positive");
    BruteFoldLists L;
    L.nfreqs   = nfreqs;
    L.seglen   = seglen;
    L.nbins    = nbins;
    L.tsamp    = tsamp;
    L.t_ref    = t_ref;
    L.use_runs = seglen > static_cast<uint32_t>(kLdsSamps);
    L.freqs.reserve(static_cast<SizeType>(nfreqs) * sizeof(double));
    hip_check(hipMemcpyAsync(L.freqs.get(), freqs_dev, nfreqs * sizeof(double),
                             hipMemcpyDeviceToDevice, s),
              "copy brute-fold frequencies");
    if (L.use_runs) {
        hip_check(hipStreamSynchronize(s), "brute-fold frequencies");
        return L;
    }
    L.counts.reserve(static_cast<SizeType>(nfreqs) * nbins * sizeof(uint32_t));
    DeviceBuffer kmax(static_cast<SizeType>(nfreqs) * sizeof(uint32_t));
    const uint32_t threads = ((nbins + 63) / 64) * 64;
    k_bf_count<<<nfreqs, threads, 0, s>>>(freqs_dev, seglen, nbins, tsamp, t_ref,
                                          L.counts.as<uint32_t>(), kmax.as<uint32_t>());
    hip_check(hipGetLastError(), "brute-fold count launch");
    std::vector<uint32_t> h_kmax(nfreqs);
    hip_check(hipMemcpyAsync(h_kmax.data(), kmax.get(), nfreqs * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, s),
              "copy bucket sizes");
    hip_check(hipStreamSynchronize(s), "brute-fold count");
    std::vector<uint64_t> h_base(nfreqs);
    uint64_t total = 0;
    for (uint32_t f = 0; f < nfreqs; ++f) {
        h_base[f] = total;
        total += static_cast<uint64_t>(h_kmax[f]) * nbins;
    }
    L.base.reserve(static_cast<SizeType>(nfreqs) * sizeof(uint64_t));
    L.list.reserve(std::max<uint64_t>(total, 1) * sizeof(uint32_t));
    hip_check(hipMemcpyAsync(L.base.get(), h_base.data(), nfreqs * sizeof(uint64_t),
                             hipMemcpyHostToDevice, s),
              "copy bucket offsets");
    k_bf_fill<<<nfreqs, threads, 0, s>>>(freqs_dev, seglen, nbins, tsamp, t_ref,
                                         L.base.as<uint64_t>(), L.list.as<uint32_t>());
    hip_check(hipGetLastError(), "brute-fold fill launch");
    // h_base must outlive its async copy.
    hip_check(hipStreamSynchronize(s), "brute-fold lists");
    return L;
}

// Device scratch of the allocation-free brute fold (brute_fold_time_noalloc):
// its frequencies, bucket counts, bucket maxima, list offsets and lists.  Each
// GPU that loads the library holds its own copy until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns on
// that device (the array's name is a host-side stand-in).
constexpr uint64_t kScratchBytes = 64ULL << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
std::array<void*, kMaxDevices> g_scratch_ptr{};

/// Device address of `scratch` on the current device, looked up once per
/// device (with that device current) and kept.
unsigned char* scratch_address() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    LOKI_REQUIRE(dev >= 0 && dev < kMaxDevices, "brute fold scratch: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[static_cast<size_t>(dev)] == nullptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                  "hipGetSymbolAddress(brute fold scratch)");
        LOKI_REQUIRE(p != nullptr, "brute fold scratch: no device address");
        g_scratch_ptr[static_cast<size_t>(dev)] = p;
    }
    return static_cast<unsigned char*>(g_scratch_ptr[static_cast<size_t>(dev)]);
}

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~static_cast<uint64_t>(255); }

/// Segments per block of the list fold: as many as fit 32 KB of LDS (8 B per
/// sample), at most 8; 1 selects the one-segment kernel.
int fold_segs_per_block(uint32_t seglen) {
    constexpr uint32_t kMsLdsSamps = 4096;
    if (seglen == 0 || seglen > kMsLdsSamps / 2) {
        return 1;
    }
    const uint32_t ns = kMsLdsSamps / seglen;
    return ns >= 8 ? 8 : (ns >= 4 ? 4 : 2);
}

template <typename TO>
void launch_fold(const float* ts_e, const float* ts_v, uint64_t blocks, uint32_t bps,
                 const BruteFoldView& L, TO* fold, float scale, hipStream_t s) {
    if (L.use_runs) {
        const auto g = static_cast<unsigned>(blocks);
        const auto* fr = L.freqs;
        k_bf_units<TO><<<g, kBlock, unit_lds_bytes(L.nbins), s>>>(
            ts_e, ts_v, L.seglen, L.nfreqs, L.nbins, bps, fr, L.tsamp, L.t_ref, fold, scale);
    } else if (fold_segs_per_block(L.seglen) > 1) {
        // blocks = nseg * bps; the multi-segment grid covers ceil(nseg / NS) groups
        const uint64_t nseg = blocks / bps;
        const int ns        = fold_segs_per_block(L.seglen);
        const uint64_t grid = (nseg + ns - 1) / ns * bps;
        const size_t lds    = static_cast<size_t>(ns) * L.seglen * sizeof(float2);
        const auto g        = static_cast<unsigned>(grid);
        const auto* li = L.list;
        const auto* ba = L.base;
        const auto* co = L.counts;
        if (ns == 8) {
            k_bf_fold_ms<8, TO><<<g, kBlock, lds, s>>>(ts_e, ts_v, L.seglen, nseg, L.nfreqs,
                                                       L.nbins, bps, li, ba, co, fold, scale);
        } else if (ns == 4) {
            k_bf_fold_ms<4, TO><<<g, kBlock, lds, s>>>(ts_e, ts_v, L.seglen, nseg, L.nfreqs,
                                                       L.nbins, bps, li, ba, co, fold, scale);
        } else {
            k_bf_fold_ms<2, TO><<<g, kBlock, lds, s>>>(ts_e, ts_v, L.seglen, nseg, L.nfreqs,
                                                       L.nbins, bps, li, ba, co, fold, scale);
        }
    } else {
        k_bf_fold<true, TO><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(
            ts_e, ts_v, L.seglen, L.nfreqs, L.nbins, bps, L.list, L.base, L.counts, fold, scale);
    }
}

}  // namespace

void brute_fold_time(const float* ts_e, const float* ts_v, uint64_t nseg,
                     const BruteFoldLists& L, void* fold, FoldStorage so, float scale,
                     hipStream_t s) {
    brute_fold_time(ts_e, ts_v, nseg, L.view(), fold, so, scale, s);
}

void brute_fold_time(const float* ts_e, const float* ts_v, uint64_t nseg,
                     const BruteFoldView& L, void* fold, FoldStorage so, float scale,
                     hipStream_t s) {
    const uint32_t bps    = brute_fold_blocks_per_segment(L);
    const uint64_t blocks = nseg * bps;
    LOKI_REQUIRE(blocks < (1ULL << 31), "brute fold grid too large");
    if (blocks == 0) {
        return;
    }
    switch (so) {
    case FoldStorage::kBFP16:
        LOKI_REQUIRE(false, "brute fold: block floating point keeps level 0 in FP32");
        return;
    case FoldStorage::kF32:
        return launch_fold(ts_e, ts_v, blocks, bps, L, static_cast<float*>(fold), scale, s);
    case FoldStorage::kF16:
        return launch_fold(ts_e, ts_v, blocks, bps, L, static_cast<_Float16*>(fold), scale, s);
    case FoldStorage::kBF16:
        return launch_fold(ts_e, ts_v, blocks, bps, L, static_cast<__bf16*>(fold), scale, s);
    }
}

uint64_t brute_fold_scratch_bytes() { return kScratchBytes; }

void brute_fold_time_noalloc(const float* ts_e, const float* ts_v, uint64_t nsamps,
                             const double* freqs_host, uint32_t nfreqs, uint32_t seglen,
                             uint32_t nbins, double tsamp, double t_ref, float* fold,
                             hipStream_t s) {
    LOKI_REQUIRE(ts_e != nullptr && ts_v != nullptr && freqs_host != nullptr && fold != nullptr,
                 "null argument");
    LOKI_REQUIRE(nfreqs >= 1, "BruteFold: Frequency array is empty");
    LOKI_REQUIRE(seglen >= 1 && nsamps % seglen == 0,
                 "BruteFold: Number of samples is not a multiple of segment length");
    LOKI_REQUIRE(nbins >= 1 && nbins <= 1024, "BruteFold: nbins must be in [1, 1024]");
    LOKI_REQUIRE(tsamp > 0.0, "brute fold: tsamp must be positive");
    // scratch layout (256-byte aligned parts): freqs f64[nfreqs] | counts
    // u32[nfreqs * nbins] | kmax u32[nfreqs] | total u64 | base u64[nfreqs] |
    // list u32[total]
    const uint64_t o_cnt   = align256(static_cast<uint64_t>(nfreqs) * sizeof(double));
    const uint64_t o_kmax  = align256(o_cnt + static_cast<uint64_t>(nfreqs) * nbins * 4);
    const uint64_t o_total = align256(o_kmax + static_cast<uint64_t>(nfreqs) * 4);
    const uint64_t o_base  = align256(o_total + sizeof(uint64_t));
    const uint64_t o_list  = align256(o_base + static_cast<uint64_t>(nfreqs) * sizeof(uint64_t));
    BruteFoldView v;
    v.nfreqs   = nfreqs;
    v.seglen   = seglen;
    v.nbins    = nbins;
    v.tsamp    = tsamp;
    v.t_ref    = t_ref;
    v.use_runs = seglen > static_cast<uint32_t>(kLdsSamps);
    const uint64_t need = v.use_runs ? static_cast<uint64_t>(nfreqs) * sizeof(double) : o_list;
    if (need > kScratchBytes) {
        throw_invalid("brute fold (no allocation): the frequencies and bucket tables need " +
                          std::to_string(need) + " bytes, more than the " +
                          std::to_string(kScratchBytes) + "-byte device scratch",
                      __FILE__, __LINE__);
    }
    unsigned char* sc = scratch_address();
    auto* d_freqs     = reinterpret_cast<double*>(sc);
    hip_check(hipMemcpyAsync(d_freqs, freqs_host, static_cast<size_t>(nfreqs) * sizeof(double),
                             hipMemcpyHostToDevice, s),
              "copy brute-fold frequencies");
    v.freqs = d_freqs;
    if (!v.use_runs) {
        auto* counts = reinterpret_cast<uint32_t*>(sc + o_cnt);
        auto* kmax   = reinterpret_cast<uint32_t*>(sc + o_kmax);
        auto* total  = reinterpret_cast<uint64_t*>(sc + o_total);
        auto* base   = reinterpret_cast<uint64_t*>(sc + o_base);
        auto* list   = reinterpret_cast<uint32_t*>(sc + o_list);
        const uint32_t threads = ((nbins + 63) / 64) * 64;
        // every count, maximum, offset and the total are written before they
        // are read; list slots past a bucket's count are never read
        k_bf_count<<<nfreqs, threads, 0, s>>>(d_freqs, seglen, nbins, tsamp, t_ref, counts, kmax);
        hip_check(hipGetLastError(), "brute-fold count launch");
        k_bf_offsets<<<1, 1024, 0, s>>>(kmax, nfreqs, nbins, base, total);
        hip_check(hipGetLastError(), "brute-fold offsets launch");
        uint64_t h_total = 0;
        hip_check(hipMemcpyAsync(&h_total, total, sizeof(uint64_t), hipMemcpyDeviceToHost, s),
                  "copy bucket total");
        hip_check(hipStreamSynchronize(s), "brute-fold count");
        const uint64_t need_list = o_list + h_total * 4;
        if (need_list > kScratchBytes) {
            throw_invalid("brute fold (no allocation): the bucket lists need " +
                              std::to_string(need_list) + " bytes, more than the " +
                              std::to_string(kScratchBytes) + "-byte device scratch",
                          __FILE__, __LINE__);
        }
        k_bf_fill<<<nfreqs, threads, 0, s>>>(d_freqs, seglen, nbins, tsamp, t_ref, base, list);
        hip_check(hipGetLastError(), "brute-fold fill launch");
        v.counts = counts;
        v.base   = base;
        v.list   = list;
    }
    brute_fold_time(ts_e, ts_v, nsamps / seglen, v, fold, FoldStorage::kF32, 1.0F, s);
    hip_check(hipGetLastError(), "brute fold launch");
    // the next call on this device reuses the scratch: finish first
    hip_check(hipStreamSynchronize(s), "brute fold");
}

uint32_t brute_fold_blocks_per_segment(const BruteFoldLists& L) {
    return brute_fold_blocks_per_segment(L.view());
}

uint32_t brute_fold_blocks_per_segment(const BruteFoldView& L) {
    if (L.use_runs) {
        const uint32_t nf = unit_nf(L.nbins);
        return (L.nfreqs + nf - 1) / nf;
    }
    const uint64_t pairs = static_cast<uint64_t>(L.nfreqs) * L.nbins;
    LOKI_REQUIRE(pairs < (1ULL << 31), "brute fold: too many (freq, bin) pairs");
    return static_cast<uint32_t>((pairs + kPairsBlk - 1) / kPairsBlk);
}

template <typename TO>
void grouped_fold(const FoldTask* tasks, uint32_t ntasks, uint32_t nblocks, bool runs,
                  uint32_t max_nbins, hipStream_t s) {
    if (!runs) {
        k_bf_fold_grouped<true, TO><<<nblocks, kBlock, 0, s>>>(tasks, ntasks);
        return;
    }
    size_t lds = 0;  // enough for every nbins up to the launch's largest
    for (uint32_t nb = 1; nb <= max_nbins; ++nb) {
        lds = std::max(lds, unit_lds_bytes(nb));
    }
    k_bf_units_grouped<TO><<<nblocks, kBlock, lds, s>>>(tasks, ntasks);
}

void brute_fold_grouped(const FoldTask* tasks, uint32_t ntasks, uint32_t nblocks, bool runs,
                        uint32_t max_nbins, FoldStorage so, hipStream_t s) {
    if (ntasks == 0 || nblocks == 0) {
        return;
    }
    switch (so) {
    case FoldStorage::kBFP16:
        LOKI_REQUIRE(false, "grouped brute fold: block floating point unsupported");
        return;
    case FoldStorage::kF32:
        return grouped_fold<float>(tasks, ntasks, nblocks, runs, max_nbins, s);
    case FoldStorage::kF16:
        return grouped_fold<_Float16>(tasks, ntasks, nblocks, runs, max_nbins, s);
    case FoldStorage::kBF16:
        return grouped_fold<__bf16>(tasks, ntasks, nblocks, runs, max_nbins, s);
    }
}

}  // namespace loki_amd::kern
