// Time-domain brute fold without atomics.
//
// The reference CPU path (fold.cpp:113-142, kernels.cpp:550-580) sums each
// (freq, phase-bin) bucket in ascending sample order; its CUDA path uses
// float atomics and is order-nondeterministic (kernels_cuda.cu:662-773).  Here
// the bucket lists are built once per plan (one block per frequency, FP64
// phases), and the fold runs one lane per (segment, freq, bin) summing its
// bucket in ascending order -- bit-identical to the CPU order and
// deterministic.  The loads of a bucket are independent of the running sum,
// so they are issued in batches and only the FP32 adds form a chain.
#include "grid.hpp"
#include "kernels.hpp"

namespace loki_amd::kern {
namespace {

constexpr int kChunk = 4096;  // samples whose bins are staged in LDS at once

// One block per frequency; thread b owns phase bin b (b < nbins).  Two passes
// over the segment: count, then (after an exclusive scan) write the sample
// indices of each bin in ascending order.
__global__ void __launch_bounds__(1024)
k_bf_buckets(const double* __restrict__ freqs, uint32_t seglen, uint32_t nbins, double tsamp,
             double t_ref, uint32_t* __restrict__ idx, uint32_t* __restrict__ off) {
    __shared__ uint16_t s_bins[kChunk];
    __shared__ uint32_t s_start[1025];
    const uint32_t f    = blockIdx.x;
    const double freq   = freqs[f];
    const uint32_t tid  = threadIdx.x;
    const uint32_t nthr = blockDim.x;

    uint32_t count = 0;
    for (uint32_t base = 0; base < seglen; base += kChunk) {
        const uint32_t len = min(static_cast<uint32_t>(kChunk), seglen - base);
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
        __syncthreads();
    }
    if (tid < nbins) {
        s_start[tid + 1] = count;
    }
    __syncthreads();
    if (tid == 0) {
        s_start[0] = 0;
        for (uint32_t b = 1; b <= nbins; ++b) {
            s_start[b] += s_start[b - 1];
        }
    }
    __syncthreads();
    uint32_t* __restrict__ off_f = off + static_cast<uint64_t>(f) * (nbins + 1);
    for (uint32_t b = tid; b <= nbins; b += nthr) {
        off_f[b] = s_start[b];
    }
    uint32_t* __restrict__ idx_f = idx + static_cast<uint64_t>(f) * seglen;
    uint32_t wpos                = (tid < nbins) ? s_start[tid] : 0U;
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
                    idx_f[wpos++] = base + k;
                }
            }
        }
    }
}

// One lane per (segment, freq, bin); consecutive lanes are consecutive bins
// of one frequency, so the two output rows of a (segment, freq) are written
// contiguously.
__global__ void __launch_bounds__(256)
k_bf_fold(const float* __restrict__ ts_e, const float* __restrict__ ts_v, uint32_t seglen,
          uint32_t nfreqs, uint32_t nbins, uint32_t blocks_per_seg,
          const uint32_t* __restrict__ idx, const uint32_t* __restrict__ off,
          float* __restrict__ fold) {
    const uint64_t seg  = blockIdx.x / blocks_per_seg;
    const uint32_t pair = (blockIdx.x % blocks_per_seg) * blockDim.x + threadIdx.x;
    if (pair >= nfreqs * nbins) {
        return;
    }
    const uint32_t f = pair / nbins;
    const uint32_t b = pair - f * nbins;
    const uint32_t* __restrict__ lst = idx + static_cast<uint64_t>(f) * seglen;
    const uint32_t beg = off[static_cast<uint64_t>(f) * (nbins + 1) + b];
    const uint32_t end = off[static_cast<uint64_t>(f) * (nbins + 1) + b + 1];
    const float* __restrict__ e = ts_e + seg * seglen;
    const float* __restrict__ v = ts_v + seg * seglen;
    float se = 0.0F;
    float sv = 0.0F;
    uint32_t i = beg;
    for (; i + 4 <= end; i += 4) {
        const uint32_t k0 = lst[i], k1 = lst[i + 1], k2 = lst[i + 2], k3 = lst[i + 3];
        const float e0 = e[k0], e1 = e[k1], e2 = e[k2], e3 = e[k3];
        const float v0 = v[k0], v1 = v[k1], v2 = v[k2], v3 = v[k3];
        se = se + e0;
        sv = sv + v0;
        se = se + e1;
        sv = sv + v1;
        se = se + e2;
        sv = sv + v2;
        se = se + e3;
        sv = sv + v3;
    }
    for (; i < end; ++i) {
        const uint32_t k = lst[i];
        se = se + e[k];
        sv = sv + v[k];
    }
    float* __restrict__ out = fold + (seg * nfreqs + f) * 2 * nbins;
    out[b]         = se;
    out[nbins + b] = sv;
}

}  // namespace

void brute_fold_buckets(const double* freqs, uint32_t nfreqs, uint32_t seglen, uint32_t nbins,
                        double tsamp, double t_ref, uint32_t* idx, uint32_t* off,
                        hipStream_t s) {
    LOKI_REQUIRE(nbins >= 1 && nbins <= 1024, "brute fold supports 1 <= nbins <= 1024");
    LOKI_REQUIRE(nfreqs >= 1 && seglen >= 1, "brute fold: empty frequency or segment");
    const uint32_t threads = ((nbins + 63) / 64) * 64;
    k_bf_buckets<<<nfreqs, threads, 0, s>>>(freqs, seglen, nbins, tsamp, t_ref, idx, off);
}

void brute_fold_time(const float* ts_e, const float* ts_v, uint64_t nseg, uint32_t seglen,
                     uint32_t nfreqs, uint32_t nbins, const uint32_t* idx, const uint32_t* off,
                     float* fold, hipStream_t s) {
    const uint64_t pairs = static_cast<uint64_t>(nfreqs) * nbins;
    LOKI_REQUIRE(pairs < (1ULL << 31), "brute fold: too many (freq, bin) pairs");
    const uint32_t bps    = static_cast<uint32_t>((pairs + 255) / 256);
    const uint64_t blocks = nseg * bps;
    LOKI_REQUIRE(blocks < (1ULL << 31), "brute fold grid too large");
    if (blocks == 0) {
        return;
    }
    k_bf_fold<<<static_cast<unsigned>(blocks), 256, 0, s>>>(ts_e, ts_v, seglen, nfreqs, nbins,
                                                            bps, idx, off, fold);
}

}  // namespace loki_amd::kern
