// Device helpers shared by the merge kernels: XCD-aware block remapping and
// the rotated float4 read of a parent row.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "common.hpp"
#include "plan.hpp"
#include "storage.hpp"

namespace loki_amd::kern {

/// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; remap
/// so XCD x works on the contiguous logical range [x*per(+rem), ...) and the
/// parent rows shared by neighbouring tiles stay inside one XCD's L2.  The
/// map is a bijection on [0, nblocks).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t kXcd = 8;
    const uint32_t per = nblocks / kXcd;
    const uint32_t rem = nblocks % kXcd;
    const uint32_t xcd = bid % kXcd;
    const uint32_t loc = bid / kXcd;
    return xcd < rem ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
}

/// Tile schedule of a merge level.  Coordinates are rows of n_f frequencies
/// (one row per (jerk, accel) cell, f fastest).  Neighbouring child rows map
/// to the same parent row (the accel grid is several times finer than its
/// parent's), so blocks are ordered
///   (segment, row group, frequency run, row within group, tile within run):
/// the `group` child rows that share parents are processed back to back for
/// one run of `run` frequency tiles, so their parent rows are served from L2
/// instead of being re-read from HBM once per child row, while each row's
/// output is still written in long sequential runs.
struct TileSchedule {
    uint64_t n_f;       ///< frequencies per row
    uint64_t n_rows;    ///< rows per segment (ncur / n_f)
    uint32_t n_ft;      ///< frequency tiles per row
    uint32_t n_runs;    ///< ceil(n_ft / run)
    uint32_t run;       ///< tiles per run
    uint32_t n_groups;  ///< ceil(n_rows / group)
    uint32_t group;     ///< rows per group
    uint32_t nblocks;   ///< nseg * n_groups * n_runs * group * run
};

struct Tile {
    uint64_t seg;
    uint64_t c0;     ///< first coordinate of the tile
    uint32_t count;  ///< coordinates in the tile (0: idle block)
};

inline TileSchedule make_schedule(uint64_t ncur, uint64_t n_f, uint64_t nseg, uint32_t tile,
                                  uint32_t group, uint32_t run) {
    LOKI_REQUIRE(n_f >= 1 && ncur % n_f == 0, "merge: ncur must be a multiple of n_freq");
    TileSchedule ts{};
    ts.n_f      = n_f;
    ts.n_rows   = ncur / n_f;
    const uint64_t n_ft = (n_f + tile - 1) / tile;
    // balanced groups of at most `group` rows (idle rows would leave whole
    // XCD ranges of the remapped grid without work)
    const uint64_t ngrp0 = (ts.n_rows + group - 1) / group;
    const uint64_t grp   = (ts.n_rows + ngrp0 - 1) / ngrp0;
    const uint64_t rn   = grp == 1 ? 1 : (n_ft < run ? n_ft : run);
    const uint64_t n_r  = (n_ft + rn - 1) / rn;
    const uint64_t n_g  = (ts.n_rows + grp - 1) / grp;
    const uint64_t nb   = nseg * n_g * n_r * grp * rn;
    LOKI_REQUIRE(nb < (1ULL << 31), "merge grid too large");
    ts.n_ft     = static_cast<uint32_t>(n_ft);
    ts.n_runs   = static_cast<uint32_t>(n_r);
    ts.run      = static_cast<uint32_t>(rn);
    ts.n_groups = static_cast<uint32_t>(n_g);
    ts.group    = static_cast<uint32_t>(grp);
    ts.nblocks  = static_cast<uint32_t>(nb);
    return ts;
}

__device__ __forceinline__ Tile tile_of(const TileSchedule& ts, uint32_t t, uint32_t tile) {
    const uint32_t per_run   = ts.group * ts.run;
    const uint32_t per_group = ts.n_runs * per_run;
    const uint32_t per_seg   = ts.n_groups * per_group;
    Tile r;
    r.seg             = t / per_seg;
    const uint32_t q  = t % per_seg;
    const uint32_t g  = q / per_group;
    const uint32_t w  = q % per_group;
    const uint32_t rb = w / per_run;
    const uint32_t w2 = w % per_run;
    const uint64_t row = static_cast<uint64_t>(g) * ts.group + (w2 / ts.run);
    const uint32_t ft  = rb * ts.run + (w2 % ts.run);
    if (row >= ts.n_rows || ft >= ts.n_ft) {
        r.c0    = 0;
        r.count = 0;
        return r;
    }
    const uint64_t f0 = static_cast<uint64_t>(ft) * tile;
    r.c0    = row * ts.n_f + f0;
    r.count = static_cast<uint32_t>(ts.n_f - f0 < tile ? ts.n_f - f0 : tile);
    return r;
}

/// Elements r..r+3 (r < 4) of the 8-element concatenation (a, b), as two
/// shift stages: by 2 when r & 2 (five selects), then by 1 when r & 1 (four).
__device__ __forceinline__ float4 funnel4(const float4 a, const float4 b, uint32_t r) {
    const bool s2 = (r & 2U) != 0;
    const bool s1 = (r & 1U) != 0;
    const float t0 = s2 ? a.z : a.x;
    const float t1 = s2 ? a.w : a.y;
    const float t2 = s2 ? b.x : a.z;
    const float t3 = s2 ? b.y : a.w;
    const float t4 = s2 ? b.z : b.x;
    float4 o;
    o.x = s1 ? t1 : t0;
    o.y = s1 ? t2 : t1;
    o.z = s1 ? t3 : t2;
    o.w = s1 ? t4 : t3;
    return o;
}

/// x mod NB for x < 2 * NB (NB a power of two).
template <int NB>
__device__ __forceinline__ uint32_t wrap_nb(uint32_t x) {
    return x & static_cast<uint32_t>(NB - 1);
}

/// kernels.cpp:78-85: integer shift u32(shift + 0.5f), nbins -> 0.
__device__ __forceinline__ uint32_t shift_int(float s, uint32_t nbins) {
    const uint32_t k = static_cast<uint32_t>(s + 0.5F);
    return k == nbins ? 0U : k;
}

/// Output bins 4*i4 .. 4*i4+3 of one component row rotated right by S
/// (out[j] = row[(j - S) mod NB]); a rotated float4 straddles at most two
/// aligned float4 chunks.
template <int NB>
__device__ __forceinline__ float4 rotated4(const float4* __restrict__ row, uint32_t i4,
                                           uint32_t S) {
    constexpr uint32_t C4 = NB / 4;
    const uint32_t e0 = (4U * i4 + NB - S) & (NB - 1);
    const uint32_t c0 = e0 >> 2;
    const uint32_t c1 = (c0 + 1) & (C4 - 1);
    const float4 a    = row[c0];
    const float4 b    = row[c1];
    return funnel4(a, b, e0 & 3U);
}

/// rotated4 for any storage type, widened to FP32: row points at the first
/// chunk of one component row of NB stored elements.  BFP16 rows are widened
/// with their scale s = 2^E (unused otherwise).
template <int NB, typename T>
__device__ __forceinline__ float4 rotated4_st(const chunk4_t<T>* __restrict__ row, uint32_t i4,
                                              uint32_t S, float s = 1.0F) {
    if constexpr (std::is_same_v<T, float>) {
        return rotated4<NB>(row, i4, S);
    } else {
        constexpr uint32_t C4 = NB / 4;
        const uint32_t e0 = (4U * i4 + NB - S) & (NB - 1);
        const uint32_t c0 = e0 >> 2;
        const uint32_t c1 = (c0 + 1) & (C4 - 1);
        const uint2 r     = funnel4_16(row[c0], row[c1], e0 & 3U);
        if constexpr (kBlockFP<T>) {
            return widen4_bfp(r, s);
        } else {
            return widen4<T>(r);
        }
    }
}

/// Aligned chunk i4 of a stored component row, widened (BFP16: scale s).
template <typename T>
__device__ __forceinline__ float4 chunk4_st(const chunk4_t<T>* __restrict__ row, uint32_t i4,
                                            float s = 1.0F) {
    if constexpr (kBlockFP<T>) {
        return widen4_bfp(row[i4], s);
    } else {
        return widen4<T>(row[i4]);
    }
}

/// Scale 2^E of component q (0: e, 1: v) of a BFP16 row pair from its
/// exponent word (low byte e, high byte v).
__device__ __forceinline__ float bfp_row_scale(const uint16_t* __restrict__ ew, uint64_t row,
                                               uint32_t q) {
    const uint32_t w = ew[row];
    return pow2f(static_cast<int>(static_cast<int8_t>(q == 0 ? (w & 0xFFU) : (w >> 8))));
}

/// Lanes per block of the time-domain merge (single-level and grouped launches).
constexpr int kMergeBlock = 256;

/// One 256-lane block of the time-domain merge (k_merge.hip): block `bid` of
/// the level's tile schedule.  TI/TO: stored types of parent and child
/// levels; when either is FP16/BF16 the FP32 sum is multiplied by `scale`
/// before it is rounded to TO.  BFP16 levels (int16_t) carry exponent planes:
/// `ein` of the parent level, `eout` of the child level (16-bit word per row
/// pair); the child's row exponents come from the row maxima of the FP32
/// sums, reduced over the C4 lanes of each component row.
/// Geometry of merge_v4_block for NB bins: lanes per row, rows per pass and
/// the passes (ITERS) whose loads are in flight together; TILE coordinates
/// per block.
template <int NB>
struct V4Geom {
    static constexpr uint32_t LPR     = NB / 2;
    static constexpr uint32_t LANES   = LPR < kMergeBlock ? LPR : kMergeBlock;
    static constexpr uint32_t ROWS    = kMergeBlock / LANES;
    static constexpr uint32_t CPL     = LPR / LANES;
    static constexpr uint32_t IT_ROWS = ROWS >= 64 ? 1U : 64U / ROWS;
    static constexpr uint32_t IT_REGS = 8U / CPL;
    static constexpr int ITERS = static_cast<int>(IT_ROWS < IT_REGS ? IT_ROWS : IT_REGS);
    static constexpr uint32_t TILE = ROWS * ITERS;
};

template <int NB, bool POLY, int ITERS, typename TI, typename TO>
__device__ __forceinline__ void merge_v4_block(uint32_t bid, const TI* __restrict__ in,
                                               TO* __restrict__ out,
                                               const FFACoord* __restrict__ coords,
                                               const FFACoordFreq* __restrict__ cf,
                                               uint64_t ncur, uint64_t nprev,
                                               const TileSchedule& ts, float scale,
                                               const uint16_t* __restrict__ ein = nullptr,
                                               uint8_t* __restrict__ eout = nullptr,
                                               bool remap = true) {
    constexpr uint32_t kBlock = kMergeBlock;
    using CI = chunk4_t<TI>;
    using CO = chunk4_t<TO>;
    constexpr uint32_t C4   = NB / 4;        // 4-element chunks per component
    constexpr uint32_t LPR  = 2 * C4;        // chunks per row
    constexpr uint32_t LANES = LPR < kBlock ? LPR : kBlock;
    constexpr uint32_t ROWS = kBlock / LANES;
    constexpr uint32_t TILE = ROWS * ITERS;
    static_assert(!(kBlockFP<TI> || kBlockFP<TO>) || (C4 <= 64 && LPR <= kBlock),
                  "block floating point: a component row must fit one wave");
    const Tile tl = tile_of(ts, remap ? xcd_remap(bid, ts.nblocks) : bid, TILE);
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
                float st = 1.0F, sh = 1.0F;
                if constexpr (kBlockFP<TI>) {
                    st = bfp_row_scale(ein, it_row, q);
                    sh = bfp_row_scale(ein, ih_row, q);
                }
                if constexpr (POLY) {
                    tv[it][k] = rotated4_st<NB, TI>(trow + q * C4, i4, St, st);
                } else {
                    tv[it][k] = chunk4_st<TI>(trow, ch, st);
                }
                hv[it][k] = rotated4_st<NB, TI>(hrow + q * C4, i4, Sh, sh);
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
                if constexpr (kScaledHalf<TI> || kScaledHalf<TO>) {
                    o.x *= scale;
                    o.y *= scale;
                    o.z *= scale;
                    o.w *= scale;
                }
                if constexpr (kBlockFP<TO>) {
                    const float mx = group_absmax<C4>(__builtin_fmaxf(
                        __builtin_fmaxf(__builtin_fabsf(o.x), __builtin_fabsf(o.y)),
                        __builtin_fmaxf(__builtin_fabsf(o.z), __builtin_fabsf(o.w))));
                    const int e = bfp_exponent(mx);
                    store_stream(&orow[lch + k * LANES], narrow4_bfp(o, pow2f(-e)));
                    const uint32_t ch = lch + k * LANES;
                    if (ch % C4 == 0) {
                        eout[(seg * ncur + c) * 2 + ch / C4] =
                            static_cast<uint8_t>(static_cast<int8_t>(e));
                    }
                } else {
                    store_stream(&orow[lch + k * LANES], narrow4<TO>(o));
                }
            }
        }
    }
}

}  // namespace loki_amd::kern
