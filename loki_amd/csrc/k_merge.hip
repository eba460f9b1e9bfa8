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

// Two merge levels in one pass (B = A merged, C = B merged): output row c of
// segment s at level C needs B's tail row T = rec_C(c).tail of segment 2s and
// head row H of segment 2s + 1, and each of those is the merge of two rows of
// A (segments 4s..4s+3) through B's shift table.  Per output bin j:
//   C[j] = round_C( round_B(A0[j - S_T - s_0] + A1[j - S_T - s_1])
//                 + round_B(A2[j - S_H - s_2] + A3[j - S_H - s_3]) )
// -- the adds and roundings of two single-level merges, in their order, so the
// bits are those of level C merged through B (kernels.cpp:71-146 twice), but
// level B never reaches HBM.  round_B / round_C are the stored type's
// rounding (identity for FP32; FP16/BF16 at scale 0.5 per level; block
// floating point with row exponents from the max over the C4 lanes of each
// component row, which together hold all NB bins of the B row).
template <typename TB>
__device__ __forceinline__ float4 round_level(float4 x, float scale) {
    if constexpr (kScaledHalf<TB>) {
        x.x = widen1(narrow1<TB>(x.x * scale));
        x.y = widen1(narrow1<TB>(x.y * scale));
        x.z = widen1(narrow1<TB>(x.z * scale));
        x.w = widen1(narrow1<TB>(x.w * scale));
    }
    return x;
}

template <int C4>
__device__ __forceinline__ float4 round_bfp_row(float4 x) {
    const float mx = group_absmax<C4>(
        __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(x.x), __builtin_fabsf(x.y)),
                        __builtin_fmaxf(__builtin_fabsf(x.z), __builtin_fabsf(x.w))));
    const int e     = bfp_exponent(mx);
    const float inv = pow2f(-e), sc = pow2f(e);
    x.x = round_bfp(x.x, inv, sc);
    x.y = round_bfp(x.y, inv, sc);
    x.z = round_bfp(x.z, inv, sc);
    x.w = round_bfp(x.w, inv, sc);
    return x;
}

template <int NB, bool POLY, int ITERS, typename TI, typename TO>
__global__ void __launch_bounds__(kBlock)
k_merge2(const TI* __restrict__ in, TO* __restrict__ out, const FFACoord* __restrict__ cb,
         const FFACoordFreq* __restrict__ cbf, const FFACoord* __restrict__ cc,
         const FFACoordFreq* __restrict__ ccf, uint64_t n_c, uint64_t n_b, uint64_t n_a,
         TileSchedule ts, float scale, const uint16_t* __restrict__ ein,
         uint8_t* __restrict__ eout) {
    using CI = chunk4_t<TI>;
    using CO = chunk4_t<TO>;
    constexpr uint32_t C4    = NB / 4;
    constexpr uint32_t LPR   = 2 * C4;
    static_assert(LPR <= static_cast<uint32_t>(kBlock) && C4 <= 64, "one lane per output chunk");
    constexpr uint32_t LANES = LPR;
    constexpr uint32_t ROWS  = kBlock / LANES;
    constexpr uint32_t TILE  = ROWS * ITERS;
    const Tile tl = tile_of(ts, xcd_remap(blockIdx.x, ts.nblocks), TILE);
    if (tl.count == 0) {
        return;
    }
    const uint64_t seg  = tl.seg;
    const uint32_t rsub = threadIdx.x / LANES;
    const uint32_t ch   = threadIdx.x % LANES;
    const uint32_t q    = ch / C4;
    const uint32_t i4   = ch % C4;
    float4 r0[ITERS], r1[ITERS], r2[ITERS], r3[ITERS];
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const uint32_t local = static_cast<uint32_t>(it) * ROWS + rsub;
        if (local < tl.count) {
            const uint64_t c = tl.c0 + local;
            uint32_t bt, bh, st, sh;  // B rows and their shifts
            uint32_t a[4], sa[4];     // A rows and their shifts within B
            if constexpr (POLY) {
                const FFACoord rc = cc[c];
                bt = rc.i_tail;
                bh = rc.i_head;
                st = shift_int(rc.shift_tail, NB);
                sh = shift_int(rc.shift_head, NB);
                const FFACoord rt = cb[bt];
                const FFACoord rh = cb[bh];
                a[0]  = rt.i_tail;
                sa[0] = shift_int(rt.shift_tail, NB);
                a[1]  = rt.i_head;
                sa[1] = shift_int(rt.shift_head, NB);
                a[2]  = rh.i_tail;
                sa[2] = shift_int(rh.shift_tail, NB);
                a[3]  = rh.i_head;
                sa[3] = shift_int(rh.shift_head, NB);
            } else {
                const FFACoordFreq rc = ccf[c];
                bt = rc.idx;
                bh = rc.idx;
                st = 0;
                sh = shift_int(rc.shift, NB);
                const FFACoordFreq rb = cbf[bt];
                a[0] = a[1] = a[2] = a[3] = rb.idx;
                sa[0] = sa[2] = 0;
                sa[1] = sa[3] = shift_int(rb.shift, NB);
            }
            float4 v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t row = (4 * seg + static_cast<uint64_t>(k)) * n_a + a[k];
                const CI* __restrict__ src = reinterpret_cast<const CI*>(in + row * (2 * NB));
                float sc = 1.0F;
                if constexpr (kBlockFP<TI>) {
                    sc = bfp_row_scale(ein, row, q);
                }
                const uint32_t rot = wrap_nb<NB>((k < 2 ? st : sh) + sa[k]);
                v[k] = rotated4_st<NB, TI>(src + q * C4, i4, rot, sc);
            }
            r0[it] = v[0];
            r1[it] = v[1];
            r2[it] = v[2];
            r3[it] = v[3];
        }
    }
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const uint32_t local = static_cast<uint32_t>(it) * ROWS + rsub;
        if (local < tl.count) {
            const uint64_t c = tl.c0 + local;
            float4 bt, bh, o;
            bt.x = r0[it].x + r1[it].x;
            bt.y = r0[it].y + r1[it].y;
            bt.z = r0[it].z + r1[it].z;
            bt.w = r0[it].w + r1[it].w;
            bh.x = r2[it].x + r3[it].x;
            bh.y = r2[it].y + r3[it].y;
            bh.z = r2[it].z + r3[it].z;
            bh.w = r2[it].w + r3[it].w;
            if constexpr (kBlockFP<TO>) {
                bt = round_bfp_row<C4>(bt);
                bh = round_bfp_row<C4>(bh);
            } else {
                bt = round_level<TO>(bt, scale);
                bh = round_level<TO>(bh, scale);
            }
            o.x = bt.x + bh.x;
            o.y = bt.y + bh.y;
            o.z = bt.z + bh.z;
            o.w = bt.w + bh.w;
            CO* __restrict__ orow = reinterpret_cast<CO*>(out + (seg * n_c + c) * (2 * NB));
            if constexpr (kBlockFP<TO>) {
                const float mx = group_absmax<C4>(__builtin_fmaxf(
                    __builtin_fmaxf(__builtin_fabsf(o.x), __builtin_fabsf(o.y)),
                    __builtin_fmaxf(__builtin_fabsf(o.z), __builtin_fabsf(o.w))));
                const int e = bfp_exponent(mx);
                store_stream(&orow[ch], narrow4_bfp(o, pow2f(-e)));
                if (i4 == 0) {
                    eout[(seg * n_c + c) * 2 + q] = static_cast<uint8_t>(static_cast<int8_t>(e));
                }
            } else {
                if constexpr (kScaledHalf<TO>) {
                    o.x *= scale;
                    o.y *= scale;
                    o.z *= scale;
                    o.w *= scale;
                }
                store_stream(&orow[ch], narrow4<TO>(o));
            }
        }
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

namespace {

template <int NB, bool POLY, typename TI, typename TO>
void launch_merge2(const TI* in, TO* out, const FFACoord* cb, const FFACoordFreq* cbf,
                   const FFACoord* cc, const FFACoordFreq* ccf, uint64_t n_c, uint64_t n_b,
                   uint64_t n_a, uint64_t nseg_c, uint64_t n_f, float scale, hipStream_t s) {
    constexpr int ITERS = 2;
    using G             = V4Geom<NB>;
    constexpr uint32_t TILE = G::ROWS * ITERS;
    const TileSchedule ts = make_schedule(n_c, n_f, nseg_c, TILE, kRowGroup, kRun);
    const uint16_t* ein = nullptr;
    uint8_t* eout       = nullptr;
    if constexpr (kBlockFP<TI>) {
        ein = reinterpret_cast<const uint16_t*>(in + (4 * nseg_c) * n_a * (2 * NB));
    }
    if constexpr (kBlockFP<TO>) {
        eout = reinterpret_cast<uint8_t*>(out + nseg_c * n_c * (2 * NB));
    }
    k_merge2<NB, POLY, ITERS, TI, TO><<<ts.nblocks, kBlock, 0, s>>>(
        in, out, cb, cbf, cc, ccf, n_c, n_b, n_a, ts, scale, ein, eout);
}

template <bool POLY, typename TI, typename TO>
void merge2_nb(const void* in, void* out, const FFACoord* cb, const FFACoordFreq* cbf,
               const FFACoord* cc, const FFACoordFreq* ccf, uint64_t n_c, uint64_t n_b,
               uint64_t n_a, uint64_t nseg_c, uint64_t n_f, uint32_t nbins, float scale,
               hipStream_t s) {
    const auto* pi = static_cast<const TI*>(in);
    auto* po       = static_cast<TO*>(out);
    switch (nbins) {
    case 8:
        return launch_merge2<8, POLY>(pi, po, cb, cbf, cc, ccf, n_c, n_b, n_a, nseg_c, n_f,
                                      scale, s);
    case 16:
        return launch_merge2<16, POLY>(pi, po, cb, cbf, cc, ccf, n_c, n_b, n_a, nseg_c, n_f,
                                       scale, s);
    case 32:
        return launch_merge2<32, POLY>(pi, po, cb, cbf, cc, ccf, n_c, n_b, n_a, nseg_c, n_f,
                                       scale, s);
    case 64:
        return launch_merge2<64, POLY>(pi, po, cb, cbf, cc, ccf, n_c, n_b, n_a, nseg_c, n_f,
                                       scale, s);
    case 128:
        return launch_merge2<128, POLY>(pi, po, cb, cbf, cc, ccf, n_c, n_b, n_a, nseg_c, n_f,
                                        scale, s);
    default: LOKI_REQUIRE(false, "two-level merge: nbins must be 8, 16, 32, 64 or 128");
    }
}

template <typename TI, typename TO>
void merge2_typed(const void* in, void* out, const FFACoord* cb, const FFACoordFreq* cbf,
                  const FFACoord* cc, const FFACoordFreq* ccf, uint64_t n_c, uint64_t n_b,
                  uint64_t n_a, uint64_t nseg_c, uint64_t n_f, uint32_t nbins, float scale,
                  hipStream_t s) {
    if (cc != nullptr) {
        merge2_nb<true, TI, TO>(in, out, cb, nullptr, cc, nullptr, n_c, n_b, n_a, nseg_c, n_f,
                                nbins, scale, s);
    } else {
        merge2_nb<false, TI, TO>(in, out, nullptr, cbf, nullptr, ccf, n_c, n_b, n_a, nseg_c,
                                 n_f, nbins, scale, s);
    }
}

}  // namespace

bool merge2_supported(uint32_t nbins) {
    return nbins == 8 || nbins == 16 || nbins == 32 || nbins == 64 || nbins == 128;
}

void merge_time2(const void* in, void* out, const FFACoord* coords_b,
                 const FFACoordFreq* coords_freq_b, const FFACoord* coords_c,
                 const FFACoordFreq* coords_freq_c, uint64_t n_c, uint64_t n_b, uint64_t n_a,
                 uint64_t nseg_c, uint64_t n_f, uint32_t nbins, FoldStorage si, FoldStorage so,
                 float scale, hipStream_t s) {
    LOKI_REQUIRE(merge2_supported(nbins), "two-level merge: unsupported nbins");
    LOKI_REQUIRE((coords_c == nullptr) == (coords_b == nullptr) &&
                     (coords_c == nullptr) != (coords_freq_c == nullptr) &&
                     (coords_freq_c == nullptr) == (coords_freq_b == nullptr),
                 "two-level merge: one kind of coordinate table for both levels");
    LOKI_REQUIRE(si == so || (si == FoldStorage::kF32 && so == FoldStorage::kBFP16),
                 "two-level merge: unsupported storage pair");
    if (n_c == 0 || nseg_c == 0) {
        return;
    }
    const auto args = std::make_tuple(in, out, coords_b, coords_freq_b, coords_c, coords_freq_c,
                                      n_c, n_b, n_a, nseg_c, n_f, nbins, scale, s);
    switch (so) {
    case FoldStorage::kF32: return std::apply(merge2_typed<float, float>, args);
    case FoldStorage::kF16: return std::apply(merge2_typed<_Float16, _Float16>, args);
    case FoldStorage::kBF16: return std::apply(merge2_typed<__bf16, __bf16>, args);
    case FoldStorage::kBFP16:
        return si == FoldStorage::kF32 ? std::apply(merge2_typed<float, int16_t>, args)
                                       : std::apply(merge2_typed<int16_t, int16_t>, args);
    }
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
