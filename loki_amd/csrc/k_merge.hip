// Time-domain FFA merge: one launch per level, no host synchronisation.
//
//   out[s, c, q, j] = in[2s, i_tail(c), q, (j - S_t) mod n]
//                   + in[2s+1, i_head(c), q, (j - S_h) mod n]
// with S = u32(shift + 0.5f), n -> 0 (kernels.cpp:71-146, 722-940;
// kernels_cuda.cu:280-383).  Freq-only levels use S_t = 0 and
// i_tail = i_head = idx.  One FP32 add per output element: bit-identical to
// the reference CPU order.
//
// Layout: a row (2 * nbins floats = e then v) is handled by nbins/2 lanes,
// each producing one float4 of output.  A rotated float4 straddles at most
// two aligned float4 chunks of the parent row, so each lane loads two
// dwordx4 per side and funnels them; the parent rows of a block's tile of
// consecutive coordinates overlap heavily and are served by L1/L2.  Blocks are
// remapped so each XCD works on a contiguous range of tiles (parent reuse
// stays inside one L2).  All offsets are 64-bit.  With FP16/BF16 storage
// (storage.hpp) a lane reads 4 elements as 8 bytes and rotates them with a
// 16-bit funnel shift; the add stays in FP32.  Block floating-point levels
// (int16 mantissas) are widened with their row's exponent and stored with an
// exponent from the row maximum of the FP32 sums (a reduction over the lanes
// of each component row).
#include <tuple>
#include <type_traits>

#include "kernels.hpp"
#include "merge_device.hpp"

namespace loki_amd::kern {
namespace {

constexpr int kBlock         = kMergeBlock;
// Child rows scheduled back to back (parent reuse in L2) and frequency tiles
// per sequential output run.  Long runs over few rows measured fastest on the
// large late levels (2 x 512: levels 11-12 of a 2^23-sample accel+freq search
// 6.34 -> 6.00 ms against 16 x 16), where the child rows outnumber the
// parent rows and the output stream dominates the traffic.
constexpr uint32_t kRowGroup = 2;
constexpr uint32_t kRun      = 512;

// TI/TO: stored element types of the parent and child levels (float, or
// _Float16 / __bf16 for reduced-precision storage).  Unless both are float the
// FP32 sum is multiplied by `scale` before it is rounded to TO (0.5 between
// reduced-precision levels, 2^(k0 + l) when the level is returned as FP32).
template <typename TI, typename TO>
constexpr bool kScaled = !(std::is_same_v<TI, float> && std::is_same_v<TO, float>);

template <int NB, bool POLY, int ITERS, typename TI, typename TO>
__global__ void __launch_bounds__(kBlock)
k_merge_v4(const TI* __restrict__ in, TO* __restrict__ out,
           const FFACoord* __restrict__ coords, const FFACoordFreq* __restrict__ cf,
           uint64_t ncur, uint64_t nprev, TileSchedule ts, float scale,
           const uint16_t* __restrict__ ein, uint8_t* __restrict__ eout) {
    merge_v4_block<NB, POLY, ITERS, TI, TO>(blockIdx.x, in, out, coords, cf, ncur, nprev, ts,
                                            scale, ein, eout);
}

// Any nbins: one lane per output element pair (e and v of bin j).
template <bool POLY, typename TI, typename TO>
__global__ void __launch_bounds__(kBlock)
k_merge_scalar(const TI* __restrict__ in, TO* __restrict__ out,
               const FFACoord* __restrict__ coords, const FFACoordFreq* __restrict__ cf,
               uint64_t ncur, uint64_t nprev, uint64_t nseg, uint32_t nbins, float scale) {
    const uint64_t total = nseg * ncur * nbins;
    for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t j   = static_cast<uint32_t>(t % nbins);
        const uint64_t r   = t / nbins;
        const uint64_t c   = r % ncur;
        const uint64_t seg = r / ncur;
        uint64_t it_row, ih_row;
        uint32_t St, Sh;
        if constexpr (POLY) {
            const FFACoord cd = coords[c];
            it_row = (2 * seg) * nprev + cd.i_tail;
            ih_row = (2 * seg + 1) * nprev + cd.i_head;
            St     = shift_int(cd.shift_tail, nbins);
            Sh     = shift_int(cd.shift_head, nbins);
        } else {
            const FFACoordFreq cd = cf[c];
            it_row = (2 * seg) * nprev + cd.idx;
            ih_row = (2 * seg + 1) * nprev + cd.idx;
            St     = 0;
            Sh     = shift_int(cd.shift, nbins);
        }
        const uint32_t jt = j < St ? j + nbins - St : j - St;
        const uint32_t jh = j < Sh ? j + nbins - Sh : j - Sh;
        const TI* __restrict__ tr = in + it_row * 2 * nbins;
        const TI* __restrict__ hr = in + ih_row * 2 * nbins;
        TO* __restrict__ o        = out + (seg * ncur + c) * 2 * nbins;
        float oe = widen1(tr[jt]) + widen1(hr[jh]);
        float ov = widen1(tr[nbins + jt]) + widen1(hr[nbins + jh]);
        if constexpr (kScaled<TI, TO>) {
            oe *= scale;
            ov *= scale;
        }
        o[j]         = narrow1<TO>(oe);
        o[nbins + j] = narrow1<TO>(ov);
    }
}

// Grouped merges of many series (one task each, sorted by block0).
template <int NB, bool POLY, int ITERS, typename T>
__global__ void __launch_bounds__(kBlock)
k_merge_grouped(const MergeTask* __restrict__ tasks, uint32_t ntasks) {
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
    const MergeTask& t = tasks[lo];
    merge_v4_block<NB, POLY, ITERS, T, T>(blockIdx.x - t.block0, static_cast<const T*>(t.in),
                                          static_cast<T*>(t.out), t.coords, t.coords_freq,
                                          t.ncur, t.nprev, t.ts, t.scale);
}

/// Rows per lane and coordinates per block of the v4 merge for NB.
template <int NB>
struct V4Geom {
    static constexpr uint32_t LPR     = NB / 2;
    static constexpr uint32_t LANES   = LPR < kBlock ? LPR : kBlock;
    static constexpr uint32_t ROWS    = kBlock / LANES;
    static constexpr uint32_t CPL     = LPR / LANES;
    static constexpr uint32_t IT_ROWS = ROWS >= 64 ? 1U : 64U / ROWS;
    static constexpr uint32_t IT_REGS = 8U / CPL;
    static constexpr int ITERS = static_cast<int>(IT_ROWS < IT_REGS ? IT_ROWS : IT_REGS);
    static constexpr uint32_t TILE = ROWS * ITERS;
};

template <int NB, bool POLY, typename T>
void launch_grouped(const MergeTask* tasks, uint32_t ntasks, uint32_t nblocks, hipStream_t s) {
    k_merge_grouped<NB, POLY, V4Geom<NB>::ITERS, T><<<nblocks, kBlock, 0, s>>>(tasks, ntasks);
}

template <bool POLY, typename T>
void grouped_nb(const MergeTask* tasks, uint32_t ntasks, uint32_t nblocks, uint32_t nbins,
                hipStream_t s) {
    switch (nbins) {
    case 8: return launch_grouped<8, POLY, T>(tasks, ntasks, nblocks, s);
    case 16: return launch_grouped<16, POLY, T>(tasks, ntasks, nblocks, s);
    case 32: return launch_grouped<32, POLY, T>(tasks, ntasks, nblocks, s);
    case 64: return launch_grouped<64, POLY, T>(tasks, ntasks, nblocks, s);
    case 128: return launch_grouped<128, POLY, T>(tasks, ntasks, nblocks, s);
    case 256: return launch_grouped<256, POLY, T>(tasks, ntasks, nblocks, s);
    default: LOKI_REQUIRE(false, "grouped merge: nbins must be a power of two in [8, 256]");
    }
}

template <int NB, bool POLY, typename TI, typename TO>
void launch_v4(const TI* in, TO* out, const FFACoord* coords, const FFACoordFreq* cf,
               uint64_t ncur, uint64_t nprev, uint64_t nseg, uint64_t n_f, float scale,
               hipStream_t s) {
    // up to 64 rows per block, at most 16 chunk loads in flight per lane
    using G               = V4Geom<NB>;
    const TileSchedule ts = make_schedule(ncur, n_f, nseg, G::TILE, kRowGroup, kRun);
    // BFP16 exponent planes follow each level's mantissas (parent level:
    // 2 * nseg segments of nprev rows)
    const uint16_t* ein = nullptr;
    uint8_t* eout       = nullptr;
    if constexpr (kBlockFP<TI>) {
        ein = reinterpret_cast<const uint16_t*>(in + (2 * nseg) * nprev * (2 * NB));
    }
    if constexpr (kBlockFP<TO>) {
        eout = reinterpret_cast<uint8_t*>(out + nseg * ncur * (2 * NB));
    }
    if constexpr ((kBlockFP<TI> || kBlockFP<TO>) && NB > 256) {
        LOKI_REQUIRE(false, "block floating-point storage needs nbins <= 256");
    } else {
        k_merge_v4<NB, POLY, G::ITERS, TI, TO><<<ts.nblocks, kBlock, 0, s>>>(
            in, out, coords, cf, ncur, nprev, ts, scale, ein, eout);
    }
}

template <bool POLY, typename TI, typename TO>
void dispatch(const TI* in, TO* out, const FFACoord* coords, const FFACoordFreq* cf,
              uint64_t ncur, uint64_t nprev, uint64_t nseg, uint64_t n_f, uint32_t nbins,
              float scale, hipStream_t s) {
    if constexpr (kBlockFP<TI> || kBlockFP<TO>) {
        LOKI_REQUIRE(nbins >= 4 && nbins <= 256 && (nbins & (nbins - 1)) == 0,
                     "block floating-point storage needs a power-of-two nbins in [4, 256]");
    }
    switch (nbins) {
    case 4: return launch_v4<4, POLY>(in, out, coords, cf, ncur, nprev, nseg, n_f, scale, s);
    case 8: return launch_v4<8, POLY>(in, out, coords, cf, ncur, nprev, nseg, n_f, scale, s);
    case 16: return launch_v4<16, POLY>(in, out, coords, cf, ncur, nprev, nseg, n_f, scale, s);
    case 32: return launch_v4<32, POLY>(in, out, coords, cf, ncur, nprev, nseg, n_f, scale, s);
    case 64: return launch_v4<64, POLY>(in, out, coords, cf, ncur, nprev, nseg, n_f, scale, s);
    case 128: return launch_v4<128, POLY>(in, out, coords, cf, ncur, nprev, nseg, n_f, scale, s);
    case 256: return launch_v4<256, POLY>(in, out, coords, cf, ncur, nprev, nseg, n_f, scale, s);
    case 512: return launch_v4<512, POLY>(in, out, coords, cf, ncur, nprev, nseg, n_f, scale, s);
    case 1024:
        return launch_v4<1024, POLY>(in, out, coords, cf, ncur, nprev, nseg, n_f, scale, s);
    default: break;
    }
    if constexpr (!(kBlockFP<TI> || kBlockFP<TO>)) {
        const uint64_t total = nseg * ncur * nbins;
        const uint64_t want  = (total + kBlock - 1) / kBlock;
        const unsigned grid  = static_cast<unsigned>(want < 65536 * 8 ? want : 65536 * 8);
        k_merge_scalar<POLY, TI, TO><<<grid, kBlock, 0, s>>>(in, out, coords, cf, ncur, nprev,
                                                             nseg, nbins, scale);
    }
}

template <typename TI, typename TO>
void merge_typed(const void* in, void* out, const FFACoord* coords,
                 const FFACoordFreq* coords_freq, uint64_t ncur, uint64_t nprev, uint64_t nseg,
                 uint64_t n_f, uint32_t nbins, float scale, hipStream_t s) {
    const auto* pi = static_cast<const TI*>(in);
    auto* po       = static_cast<TO*>(out);
    if (coords != nullptr) {
        dispatch<true>(pi, po, coords, nullptr, ncur, nprev, nseg, n_f, nbins, scale, s);
    } else {
        dispatch<false>(pi, po, nullptr, coords_freq, ncur, nprev, nseg, n_f, nbins, scale, s);
    }
}

}  // namespace

void merge_time(const void* in, void* out, const FFACoord* coords,
                const FFACoordFreq* coords_freq, uint64_t ncur, uint64_t nprev, uint64_t nseg,
                uint64_t n_f, uint32_t nbins, FoldStorage si, FoldStorage so, float scale,
                hipStream_t s) {
    LOKI_REQUIRE(nbins >= 1, "merge: nbins must be positive");
    LOKI_REQUIRE((coords == nullptr) != (coords_freq == nullptr),
                 "merge: exactly one coordinate table");
    LOKI_REQUIRE(si == so || so == FoldStorage::kF32 ||
                     (si == FoldStorage::kF32 && so == FoldStorage::kBFP16),
                 "merge: unsupported storage pair");
    if (ncur == 0 || nseg == 0) {
        return;
    }
    const auto args = std::make_tuple(in, out, coords, coords_freq, ncur, nprev, nseg, n_f, nbins,
                                      scale, s);
    if (si == FoldStorage::kF32) {
        return so == FoldStorage::kBFP16 ? std::apply(merge_typed<float, int16_t>, args)
                                         : std::apply(merge_typed<float, float>, args);
    }
    if (si == FoldStorage::kBFP16) {
        return so == FoldStorage::kBFP16 ? std::apply(merge_typed<int16_t, int16_t>, args)
                                         : std::apply(merge_typed<int16_t, float>, args);
    }
    if (si == FoldStorage::kF16) {
        return so == FoldStorage::kF16 ? std::apply(merge_typed<_Float16, _Float16>, args)
                                       : std::apply(merge_typed<_Float16, float>, args);
    }
    return so == FoldStorage::kBF16 ? std::apply(merge_typed<__bf16, __bf16>, args)
                                    : std::apply(merge_typed<__bf16, float>, args);
}

uint32_t merge_tile(uint32_t nbins) {
    switch (nbins) {
    case 8: return V4Geom<8>::TILE;
    case 16: return V4Geom<16>::TILE;
    case 32: return V4Geom<32>::TILE;
    case 64: return V4Geom<64>::TILE;
    case 128: return V4Geom<128>::TILE;
    case 256: return V4Geom<256>::TILE;
    default: LOKI_REQUIRE(false, "grouped merge: nbins must be a power of two in [8, 256]");
    }
    return 0;
}

void merge_grouped(const MergeTask* tasks, uint32_t ntasks, uint32_t nblocks, uint32_t nbins,
                   bool poly, FoldStorage st, hipStream_t s) {
    if (ntasks == 0 || nblocks == 0) {
        return;
    }
    LOKI_REQUIRE(st != FoldStorage::kBFP16, "grouped merge: block floating point unsupported");
    if (st == FoldStorage::kF32) {
        return poly ? grouped_nb<true, float>(tasks, ntasks, nblocks, nbins, s)
                    : grouped_nb<false, float>(tasks, ntasks, nblocks, nbins, s);
    }
    if (st == FoldStorage::kF16) {
        return poly ? grouped_nb<true, _Float16>(tasks, ntasks, nblocks, nbins, s)
                    : grouped_nb<false, _Float16>(tasks, ntasks, nblocks, nbins, s);
    }
    return poly ? grouped_nb<true, __bf16>(tasks, ntasks, nblocks, nbins, s)
                : grouped_nb<false, __bf16>(tasks, ntasks, nblocks, nbins, s);
}

}  // namespace loki_amd::kern
