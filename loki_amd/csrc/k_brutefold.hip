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
#include "fold_device.hpp"
#include "kernels.hpp"

namespace loki_amd::kern {
namespace {

constexpr int kChunk = 4096;  // samples whose bins are counted in LDS at once
constexpr int kBlock = kFoldBlock;

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
    __shared__ float s_e[USE_LDS ? kLdsSamps : 1];
    __shared__ float s_v[USE_LDS ? kLdsSamps : 1];
    bf_fold_block<USE_LDS, TO>(blockIdx.x, ts_e, ts_v, seglen, nfreqs, nbins, blocks_per_seg,
                               list, base_f, counts, fold, scale, s_e, s_v);
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
    __shared__ float s_e[USE_LDS ? kLdsSamps : 1];
    __shared__ float s_v[USE_LDS ? kLdsSamps : 1];
    bf_fold_block<USE_LDS, TO>(blockIdx.x - t.block0, t.ts_e, t.ts_v, t.seglen, t.nfreqs,
                               t.nbins, t.bps, t.list, t.base, t.counts,
                               static_cast<TO*>(t.fold), t.scale, s_e, s_v);
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
