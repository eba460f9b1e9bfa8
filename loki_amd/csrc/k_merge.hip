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
// 16-bit funnel shift; the add stays in FP32.
#include <tuple>
#include <type_traits>

#include "kernels.hpp"
#include "merge_device.hpp"

namespace loki_amd::kern {
namespace {

constexpr int kBlock         = 256;
constexpr uint32_t kRowGroup = 16;  // child rows scheduled back to back (parent reuse in L2)
constexpr uint32_t kRun      = 16;  // frequency tiles per sequential output run

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
           uint64_t ncur, uint64_t nprev, TileSchedule ts, float scale) {
    using CI = chunk4_t<TI>;
    using CO = chunk4_t<TO>;
    constexpr uint32_t C4   = NB / 4;        // 4-element chunks per component
    constexpr uint32_t LPR  = 2 * C4;        // chunks per row
    constexpr uint32_t LANES = LPR < kBlock ? LPR : kBlock;
    constexpr uint32_t ROWS = kBlock / LANES;
    constexpr uint32_t TILE = ROWS * ITERS;
    const Tile tl = tile_of(ts, xcd_remap(blockIdx.x, ts.nblocks), TILE);
    if (tl.count == 0) {
        return;
    }
    const uint64_t seg  = tl.seg;
    const uint32_t rsub = threadIdx.x / LANES;
    const uint32_t lch  = threadIdx.x % LANES;
    const uint64_t in_seg_t = (2 * seg) * nprev;
    const uint64_t in_seg_h = (2 * seg + 1) * nprev;
    constexpr uint32_t CPL = LPR / LANES;  // chunks per lane per row
    // All loads of the tile are issued before any add so several rows are in
    // flight per lane.
    float4 tv[ITERS][CPL], hv[ITERS][CPL];
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const uint32_t local = static_cast<uint32_t>(it) * ROWS + rsub;
        const uint64_t c     = tl.c0 + local;
        if (local < tl.count) {
            uint64_t it_row, ih_row;
            uint32_t St, Sh;
            if constexpr (POLY) {
                const FFACoord cd = coords[c];
                it_row = in_seg_t + cd.i_tail;
                ih_row = in_seg_h + cd.i_head;
                St     = shift_int(cd.shift_tail, NB);
                Sh     = shift_int(cd.shift_head, NB);
            } else {
                const FFACoordFreq cd = cf[c];
                it_row = in_seg_t + cd.idx;
                ih_row = in_seg_h + cd.idx;
                St     = 0;
                Sh     = shift_int(cd.shift, NB);
            }
            const CI* __restrict__ trow = reinterpret_cast<const CI*>(in + it_row * (2 * NB));
            const CI* __restrict__ hrow = reinterpret_cast<const CI*>(in + ih_row * (2 * NB));
#pragma unroll
            for (uint32_t k = 0; k < CPL; ++k) {
                const uint32_t ch = lch + k * LANES;
                const uint32_t q  = ch / C4;
                const uint32_t i4 = ch % C4;
                if constexpr (POLY) {
                    tv[it][k] = rotated4_st<NB, TI>(trow + q * C4, i4, St);
                } else {
                    tv[it][k] = widen4<TI>(trow[ch]);
                }
                hv[it][k] = rotated4_st<NB, TI>(hrow + q * C4, i4, Sh);
            }
        }
    }
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const uint32_t local = static_cast<uint32_t>(it) * ROWS + rsub;
        const uint64_t c     = tl.c0 + local;
        if (local < tl.count) {
            CO* __restrict__ orow = reinterpret_cast<CO*>(out + (seg * ncur + c) * (2 * NB));
#pragma unroll
            for (uint32_t k = 0; k < CPL; ++k) {
                float4 o;
                o.x = tv[it][k].x + hv[it][k].x;
                o.y = tv[it][k].y + hv[it][k].y;
                o.z = tv[it][k].z + hv[it][k].z;
                o.w = tv[it][k].w + hv[it][k].w;
                if constexpr (kScaled<TI, TO>) {
                    o.x *= scale;
                    o.y *= scale;
                    o.z *= scale;
                    o.w *= scale;
                }
                orow[lch + k * LANES] = narrow4<TO>(o);
            }
        }
    }
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

template <int NB, bool POLY, typename TI, typename TO>
void launch_v4(const TI* in, TO* out, const FFACoord* coords, const FFACoordFreq* cf,
               uint64_t ncur, uint64_t nprev, uint64_t nseg, uint64_t n_f, float scale,
               hipStream_t s) {
    constexpr uint32_t LPR   = NB / 2;
    constexpr uint32_t LANES = LPR < kBlock ? LPR : kBlock;
    constexpr uint32_t ROWS  = kBlock / LANES;
    constexpr uint32_t CPL   = LPR / LANES;
    // rows per lane: up to 64 rows per block, at most 16 chunk loads in flight per lane
    constexpr uint32_t IT_ROWS = ROWS >= 64 ? 1U : 64U / ROWS;
    constexpr uint32_t IT_REGS = 8U / CPL;
    constexpr int ITERS      = static_cast<int>(IT_ROWS < IT_REGS ? IT_ROWS : IT_REGS);
    constexpr uint32_t TILE  = ROWS * ITERS;
    const TileSchedule ts    = make_schedule(ncur, n_f, nseg, TILE, kRowGroup, kRun);
    k_merge_v4<NB, POLY, ITERS, TI, TO><<<ts.nblocks, kBlock, 0, s>>>(in, out, coords, cf, ncur,
                                                                      nprev, ts, scale);
}

template <bool POLY, typename TI, typename TO>
void dispatch(const TI* in, TO* out, const FFACoord* coords, const FFACoordFreq* cf,
              uint64_t ncur, uint64_t nprev, uint64_t nseg, uint64_t n_f, uint32_t nbins,
              float scale, hipStream_t s) {
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
    const uint64_t total = nseg * ncur * nbins;
    const uint64_t want  = (total + kBlock - 1) / kBlock;
    const unsigned grid  = static_cast<unsigned>(want < 65536 * 8 ? want : 65536 * 8);
    k_merge_scalar<POLY, TI, TO><<<grid, kBlock, 0, s>>>(in, out, coords, cf, ncur, nprev, nseg,
                                                         nbins, scale);
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
    LOKI_REQUIRE(si == so || so == FoldStorage::kF32, "merge: unsupported storage pair");
    if (ncur == 0 || nseg == 0) {
        return;
    }
    const auto args = std::make_tuple(in, out, coords, coords_freq, ncur, nprev, nseg, n_f, nbins,
                                      scale, s);
    if (si == FoldStorage::kF32) {
        return std::apply(merge_typed<float, float>, args);
    }
    if (si == FoldStorage::kF16) {
        return so == FoldStorage::kF16 ? std::apply(merge_typed<_Float16, _Float16>, args)
                                       : std::apply(merge_typed<_Float16, float>, args);
    }
    return so == FoldStorage::kBF16 ? std::apply(merge_typed<__bf16, __bf16>, args)
                                    : std::apply(merge_typed<__bf16, float>, args);
}

}  // namespace loki_amd::kern
