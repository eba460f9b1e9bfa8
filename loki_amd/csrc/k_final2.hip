// Fused last FFA levels, one lane per trial: FP64 shift tables + merge + boxcar
// S/N, with the parent rows staged once per block in LDS.
//
// compute_ffa_scores (ffa.cpp:444-466) writes the final fold cube and reads
// it back for snr_boxcar_3d.  Here each block owns a chunk of final grid rows
// (flattened (jerk, accel) cells) that share their parent rows, over a run of
// consecutive frequencies; one wave covers 64 frequencies of one row and each
// lane one trial:
//   0. the lane resolves its trial in FP64 (taylor_ffa.cpp:19-271 order, the
//      same bits as the table kernels): parent indices and shifts, from the
//      grid row's frequency-independent terms formed once per row on the host;
//   1. the block's parent rows -- a contiguous range per side, since the
//      chunk's rows share parent rows and the parent frequency index is
//      monotone in the trial frequency -- are staged in LDS as (e, v) pairs
//      per bin, with bins 0..2 repeated after the last bin so that any four
//      consecutive bins of a rotated row are one contiguous read.  The rows
//      are either copied from the stored parent level or, with `grand` set,
//      merged on chip from the grandparent level through the parent level's
//      shift table (kernels.cpp:71-146: one FP32 add per element, then the
//      reduced-precision rounding of a stored level -- for block floating
//      point the row exponent comes from a max over the row's lanes), so the
//      parent level never reaches HBM;
//   2. the lane merges its tail and head rows (one FP32 add each), forms
//      p = e * (1/sqrt(v)) (score.cpp:139-141) and the sequential circular
//      prefix sum in registers;
//   3. it scores every width from register windows (score.cpp:95-172):
//      snr = (h + b) * dmax - b * sum.
// Every FP32 operation matches the unfused path and the CPU oracle bit for
// bit.  The LDS allocation holds the widest parent range of any block of
// the launch, found beforehand by final2_ranges with the same resolver.
#include <cfloat>
#include <type_traits>
#include <utility>

#include "merge_device.hpp"
#include "resolve_device.hpp"
#include "score_device.hpp"

namespace loki_amd::kern {
namespace {

constexpr uint32_t kLanes = 64;
constexpr uint32_t kWaves = 4;
constexpr uint32_t kBlock = kWaves * kLanes;
/// Grid rows per block (rows sharing their parent rows): wave w takes rows
/// w, w + 4, w + 8 and w + 12 of its chunk, so the staged parent rows serve
/// up to 16 rows of trials.  Chunks of at most 4 rows run the one-row-per-wave
/// instance (RPW = 1).
constexpr uint32_t kRowsPerWave = 4;
constexpr uint32_t kMaxRows     = kWaves * kRowsPerWave;

/// Repeated leading bins of a staged row: a rotated read of 4 bins starting
/// at any bin stays inside the row.
constexpr uint32_t kPad = 3;

/// (e, v) pairs per staged row: at least NB + kPad, and 5 mod 16.  Lane l of
/// a wave reads bin k of its rotated parent row, at pair r_l * stride +
/// (k - S_l) mod NB; along a grid row the parent index r_l and the shift S_l
/// both advance by about half a step per trial at the FFA-native spacing
/// (more shift per trial for coarser spacings), so lanes whose rows differ
/// by d and shifts by about d, 2d, 3d or 4d must land on different LDS bank
/// pairs: stride * d - m * d != 0 mod 16 for m = 0..4.  (A stride of 1 mod 16
/// makes every lane of such a wave collide: 45 % of the kernel's cycles were
/// bank conflicts.)
template <int NB>
constexpr uint32_t kStride = ((NB + kPad + 10) / 16) * 16 + 5;
static_assert(kStride<32> == 37 && kStride<16> == 21 && kStride<8> == 21);

/// Reduced-precision rounding of a stored level: FP32 -> TI -> FP32 (exact
/// round trip of the rounded value); identity for FP32.
template <typename TI>
__device__ __forceinline__ float round_stored(float x) {
    if constexpr (std::is_same_v<TI, float>) {
        return x;
    } else {
        return widen1(narrow1<TI>(x));
    }
}

template <int NB>
__device__ __forceinline__ uint32_t wrap(uint32_t b) {
    return b & static_cast<uint32_t>(NB - 1);
}

/// Parent record of level (F-1) coordinate k for the grandparent merge.
template <int NP>
__device__ __forceinline__ void parent_record(const Final2Args& a, uint32_t k, uint32_t& it,
                                              uint32_t& st, uint32_t& ih, uint32_t& sh,
                                              uint32_t nb) {
    if constexpr (NP == 1) {
        const FFACoordFreq r = a.coords1f[k];
        it = r.idx;
        ih = r.idx;
        st = 0;
        sh = shift_int(r.shift, nb);
    } else {
        const FFACoord r = a.coords1[k];
        it = r.i_tail;
        ih = r.i_head;
        st = shift_int(r.shift_tail, nb);
        sh = shift_int(r.shift_head, nb);
    }
}

/// max_j (hi_j - ps[j]) for the fixed width W, hi_j = ps[j + W] or, past the
/// end of the row, ps[j + W - NB] + sum (the circular extension of
/// score.cpp:143-156), formed where it is used: holding the NB - 1 extended
/// values across the widths costs 31 VGPRs (96 instead of 71, 5 waves per
/// SIMD instead of 7) for 45 fewer adds per trial.
template <int NB, int W>
__device__ __forceinline__ float window_max_fixed(const float (&ps)[NB]) {
    float ma = -FLT_MAX, mb = -FLT_MAX;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const float hi = (j + W < NB) ? ps[(j + W) % NB] : ps[(j + W - NB) % NB] + ps[NB - 1];
        const float d  = hi - ps[j];
        if ((j & 2) == 0) {
            ma = __builtin_fmaxf(ma, d);
        } else {
            mb = __builtin_fmaxf(mb, d);
        }
    }
    return __builtin_fmaxf(ma, mb);
}

template <int NB, int... Ws>
__device__ __forceinline__ float window_max_reg(std::integer_sequence<int, Ws...>,
                                                const float (&ps)[NB], uint32_t w) {
    float r = 0.0F;
    (void)((w == static_cast<uint32_t>(Ws + 1) ? (r = window_max_fixed<NB, Ws + 1>(ps), true)
                                               : false) ||
           ...);
    return r;
}

template <bool EXACT>
__device__ __forceinline__ float4 normalise4v(const float4 e, const float4 v) {
    if constexpr (EXACT) {
        return normalise4_exact(e, v);
    } else {
        float4 p;
        p.x = normalise<false>(e.x, v.x);
        p.y = normalise<false>(e.y, v.y);
        p.z = normalise<false>(e.z, v.z);
        p.w = normalise<false>(e.w, v.w);
        return p;
    }
}

template <int NB, int NP, int RPW, bool FINDMAX, bool EXACT, typename TI>
__global__ void __launch_bounds__(kBlock)
k_final2(Final2Args a) {
    constexpr uint32_t C4     = NB / 4;           // 4-bin chunks per row
    constexpr uint32_t STRIDE = kStride<NB>;      // (e, v) pairs per staged row
    extern __shared__ float2 s_rows[];            // [2 * cap][STRIDE]
    __shared__ uint32_t s_rng[4];                 // lo_t, hi_t, lo_h, hi_h

    // rows wave, wave + 4, ... of the chunk (wave-uniform: the row terms are scalar loads)
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
    const uint32_t lane = threadIdx.x % kLanes;  // frequency of the tile
    const uint32_t nthr = blockDim.x;
    const uint32_t id   = xcd_remap(blockIdx.x, gridDim.x);
    // frequency-tile-major over the chunks: concurrently running blocks cover
    // every row chunk at neighbouring frequencies, so the grandparent rows
    // they share stay in L2
    const uint32_t tile  = id / a.n_chunks;
    const uint32_t chunk = id % a.n_chunks;
    const uint32_t row0  = a.chunks[chunk].row0;
    const uint32_t nrows = a.chunks[chunk].nrows;
    const uint64_t jf64  = static_cast<uint64_t>(tile) * kLanes + lane;
    const uint32_t n_f   = a.lv.n_cur[2];
    const bool in_tile   = jf64 < n_f;
    const uint32_t jf    = static_cast<uint32_t>(jf64);
    auto has_row = [&](uint32_t k) { return in_tile && wave + k * kWaves < nrows; };

    // 0. this lane's trials (one per row of the wave)
    FFACoord cds[RPW];
    uint32_t lt = 0xFFFFFFFFU, ht = 0, lh = 0xFFFFFFFFU, hh = 0;
#pragma unroll
    for (uint32_t k = 0; k < RPW; ++k) {
        cds[k] = FFACoord{0, 0.0F, 0, 0.0F};
        if (has_row(k)) {
            if constexpr (NP == 1) {
                cds[k] = resolve_pair_rc<1>(a.lv, a.fg.dv, row0 + wave + k * kWaves, jf);
            } else {
                // the row's terms were formed once per grid row (same operations)
                const RowTerms* rt = a.rows + 2 * static_cast<size_t>(row0 + wave + k * kWaves);
                cds[k]             = resolve_pair_rt(a.lv, a.fg.dv, rt[0], rt[1], jf);
            }
            lt     = min(lt, cds[k].i_tail);
            ht     = max(ht, cds[k].i_tail);
            lh     = min(lh, cds[k].i_head);
            hh     = max(hh, cds[k].i_head);
        }
    }
    if (threadIdx.x < 4) {
        s_rng[threadIdx.x] = (threadIdx.x % 2 == 0) ? 0xFFFFFFFFU : 0U;
    }
    __syncthreads();
    {
        // the parent indices are non-decreasing in the frequency along a grid
        // row (monotone roundings of f * (1 - v/c)), so the wave's minima sit
        // in lane 0 and its maxima in the last lane inside the grid
        const uint64_t f0   = static_cast<uint64_t>(tile) * kLanes;
        const uint32_t last = f0 < n_f ? static_cast<uint32_t>(
                                             min<uint64_t>(kLanes - 1, n_f - 1 - f0))
                                       : 0U;
        lt = __builtin_amdgcn_readfirstlane(lt);
        lh = __builtin_amdgcn_readfirstlane(lh);
        ht = __builtin_amdgcn_readlane(ht, last);
        hh = __builtin_amdgcn_readlane(hh, last);
        if (lane == 0 && lt != 0xFFFFFFFFU) {
            atomicMin(&s_rng[0], lt);
            atomicMax(&s_rng[1], ht);
            atomicMin(&s_rng[2], lh);
            atomicMax(&s_rng[3], hh);
        }
    }
    __syncthreads();
    const uint32_t lo_t = s_rng[0];
    const uint32_t lo_h = s_rng[2];
    if (lo_t == 0xFFFFFFFFU) {
        return;  // no trial in this block (uniform: every wave returns)
    }
    const uint32_t n_t = s_rng[1] - lo_t + 1;
    const uint32_t n_h = s_rng[3] - lo_h + 1;
    // each trial's parent rows as staged slots and integer shifts, one word
    // per row (slots < 256: the LDS budget bounds cap; shifts < NB)
    uint32_t trial[RPW];
    bool outside = n_t > a.cap || n_h > a.cap;
#pragma unroll
    for (uint32_t k = 0; k < RPW; ++k) {
        const uint32_t St = NP == 1 ? 0U : shift_int(cds[k].shift_tail, NB);
        const uint32_t Sh = shift_int(cds[k].shift_head, NB);
        const uint32_t dt = cds[k].i_tail - lo_t;
        const uint32_t dh = cds[k].i_head - lo_h;
        outside           = outside || (has_row(k) && (dt >= n_t || dh >= n_h));
        trial[k]          = (dt & 0xFFU) | (St << 8) | ((dh & 0xFFU) << 16) | (Sh << 24);
    }
    if (__syncthreads_or(outside ? 1 : 0) != 0) {
        // cannot happen: cap is the exact maximum over all blocks
        // (final2_ranges) and the range holds every trial of the block; never
        // index past the staged rows
        for (uint32_t k = 0; k < RPW; ++k) {
            if (has_row(k) && a.fo.count == nullptr) {
                const uint64_t c_idx =
                    static_cast<uint64_t>(row0 + wave + k * kWaves) * n_f + jf;
                const uint32_t nw = FINDMAX ? 1U : a.wt.n;
                for (uint32_t iw = 0; iw < nw; ++iw) {
                    a.scores[c_idx * nw + iw] = __builtin_nanf("");
                }
            }
        }
        return;
    }

    // 1. stage the parent rows: task (row slot, CH chunks qp, qp + TPS, ... of
    //    4 bins); `unsafe` marks a staged variance outside [2^-96, 2^126), the
    //    range whose pairwise sums the short exact normalisation takes
    constexpr uint32_t CH  = 1;         // 4-bin chunks per task (2: no faster)
    constexpr uint32_t TPS = C4 / CH;   // tasks per staged row
    bool unsafe = !EXACT || kScaledHalf<TI>;  // scaled storage: sums are rescaled first
    {
        const uint32_t ntask = (n_t + n_h) * TPS;
        for (uint32_t t = threadIdx.x; t < ntask; t += nthr) {
            const uint32_t slot = t / TPS;
            const uint32_t qp   = t % TPS;
            const uint32_t side = slot < n_t ? 0U : 1U;
            const uint32_t k    = side == 0 ? lo_t + slot : lo_h + (slot - n_t);
            float2* dst         = s_rows + (side * a.cap + (side == 0 ? slot : slot - n_t)) * STRIDE;
            using CI            = chunk4_t<TI>;
            const TI* in        = static_cast<const TI*>(a.in);
            float4 e[CH], v[CH];
            uint32_t b0 = 0;  // sums of chunk q4 go to bins 4*q4 + i + b0
            if (!a.grand) {
                const uint64_t r = (side * a.n_p1) + k;
                const CI* src    = reinterpret_cast<const CI*>(in + r * (2 * NB));
                float se = 1.0F, sv = 1.0F;
                if constexpr (kBlockFP<TI>) {
                    se = bfp_row_scale(a.ein, r, 0);
                    sv = bfp_row_scale(a.ein, r, 1);
                }
#pragma unroll
                for (uint32_t h = 0; h < CH; ++h) {
                    e[h] = chunk4_st<TI>(src, qp + h * TPS, se);
                    v[h] = chunk4_st<TI>(src + C4, qp + h * TPS, sv);
                }
            } else {
                uint32_t it, st, ih, sh;
                parent_record<NP>(a, k, it, st, ih, sh, NB);
                // rows (2 side) and (2 side + 1) of the grandparent level: a
                // side's tail half selected once, each row one 64-bit multiply-add
                const uint64_t half = a.n_p2 * (2 * NB);
                const TI* tbase     = in + (side != 0 ? 2 * half : 0);
                const CI* trow =
                    reinterpret_cast<const CI*>(tbase + static_cast<uint64_t>(it) * (2 * NB));
                const CI* hrow = reinterpret_cast<const CI*>(
                    tbase + half + static_cast<uint64_t>(ih) * (2 * NB));
                float ste = 1.0F, stv = 1.0F, she = 1.0F, shv = 1.0F;
                if constexpr (kBlockFP<TI>) {
                    const uint64_t rt = ((2 * side) * a.n_p2) + it;
                    const uint64_t rh = ((2 * side + 1) * a.n_p2) + ih;
                    ste = bfp_row_scale(a.ein, rt, 0);
                    stv = bfp_row_scale(a.ein, rt, 1);
                    she = bfp_row_scale(a.ein, rh, 0);
                    shv = bfp_row_scale(a.ein, rh, 1);
                }
                // parent[j] = tail[j - St] + head[j - Sh]: the tail is read
                // unrotated and the sums placed at j = 4*q4 + i + St, so only
                // the head is rotated (by Sh - St)
                const uint32_t d = wrap<NB>(sh + NB - st);
#pragma unroll
                for (uint32_t h = 0; h < CH; ++h) {
                    const uint32_t q4 = qp + h * TPS;
                    const float4 te   = chunk4_st<TI>(trow, q4, ste);
                    const float4 tv   = chunk4_st<TI>(trow + C4, q4, stv);
                    const float4 he   = rotated4_st<NB, TI>(hrow, q4, d, she);
                    const float4 hv   = rotated4_st<NB, TI>(hrow + C4, q4, d, shv);
                    e[h].x = te.x + he.x;
                    e[h].y = te.y + he.y;
                    e[h].z = te.z + he.z;
                    e[h].w = te.w + he.w;
                    v[h].x = tv.x + hv.x;
                    v[h].y = tv.y + hv.y;
                    v[h].z = tv.z + hv.z;
                    v[h].w = tv.w + hv.w;
                    // one chunk's loads in flight at a time (with CH = 2: 60
                    // VGPRs, not 76)
                    __builtin_amdgcn_sched_barrier(0);
                }
                bool bfp_parent = kBlockFP<TI>;
                if constexpr (std::is_same_v<TI, float>) {
                    bfp_parent = a.pbfp != 0;  // FP32 brute fold under a BFP16 parent
                }
                if (bfp_parent) {
                    // the parent level as stored: row exponents from the row
                    // maxima over the TPS consecutive tasks (lanes) of this row
                    float me = 0.0F, mv = 0.0F;
#pragma unroll
                    for (uint32_t h = 0; h < CH; ++h) {
                        me = __builtin_fmaxf(
                            me, __builtin_fmaxf(
                                    __builtin_fmaxf(__builtin_fabsf(e[h].x), __builtin_fabsf(e[h].y)),
                                    __builtin_fmaxf(__builtin_fabsf(e[h].z), __builtin_fabsf(e[h].w))));
                        mv = __builtin_fmaxf(
                            mv, __builtin_fmaxf(
                                    __builtin_fmaxf(__builtin_fabsf(v[h].x), __builtin_fabsf(v[h].y)),
                                    __builtin_fmaxf(__builtin_fabsf(v[h].z), __builtin_fabsf(v[h].w))));
                    }
                    const int ee = bfp_exponent(group_absmax<TPS>(me));
                    const int ev = bfp_exponent(group_absmax<TPS>(mv));
                    const float ie = pow2f(-ee), se = pow2f(ee);
                    const float iv = pow2f(-ev), sv = pow2f(ev);
#pragma unroll
                    for (uint32_t h = 0; h < CH; ++h) {
                        e[h].x = round_bfp(e[h].x, ie, se);
                        e[h].y = round_bfp(e[h].y, ie, se);
                        e[h].z = round_bfp(e[h].z, ie, se);
                        e[h].w = round_bfp(e[h].w, ie, se);
                        v[h].x = round_bfp(v[h].x, iv, sv);
                        v[h].y = round_bfp(v[h].y, iv, sv);
                        v[h].z = round_bfp(v[h].z, iv, sv);
                        v[h].w = round_bfp(v[h].w, iv, sv);
                    }
                } else if constexpr (kScaledHalf<TI>) {
#pragma unroll
                    for (uint32_t h = 0; h < CH; ++h) {
                        e[h].x = round_stored<TI>(e[h].x * a.scale1);
                        e[h].y = round_stored<TI>(e[h].y * a.scale1);
                        e[h].z = round_stored<TI>(e[h].z * a.scale1);
                        e[h].w = round_stored<TI>(e[h].w * a.scale1);
                        v[h].x = round_stored<TI>(v[h].x * a.scale1);
                        v[h].y = round_stored<TI>(v[h].y * a.scale1);
                        v[h].z = round_stored<TI>(v[h].z * a.scale1);
                        v[h].w = round_stored<TI>(v[h].w * a.scale1);
                    }
                }
                b0 = st;
            }
#pragma unroll
            for (uint32_t h = 0; h < CH; ++h) {
                const float ev[4][2] = {{e[h].x, v[h].x}, {e[h].y, v[h].y}, {e[h].z, v[h].z},
                                        {e[h].w, v[h].w}};
                const uint32_t bq    = 4 * (qp + h * TPS) + b0;
#pragma unroll
                for (uint32_t i = 0; i < 4; ++i) {
                    const uint32_t b = wrap<NB>(bq + i);
                    const float2 pr  = make_float2(ev[i][0], ev[i][1]);
                    dst[b]           = pr;
                    if (b < kPad) {
                        dst[NB + b] = pr;
                    }
                }
                if constexpr (!EXACT) {
                    continue;
                }
                const uint32_t lo = 0x0F800000U;  // 2^-96
                const uint32_t hi = 0x7E800000U;  // 2^126
                const uint32_t m  = max(max(__float_as_uint(v[h].x) - lo, __float_as_uint(v[h].y) - lo),
                                        max(__float_as_uint(v[h].z) - lo, __float_as_uint(v[h].w) - lo));
                unsafe            = unsafe || m >= hi - lo;
            }
        }
    }
    // every trial of the block sums two staged variances: with all of them in
    // [2^-96, 2^126) each sum is in [2^-96, 2^127), inside the short sequences'
    // range, and the per-chunk range checks are skipped (block-uniform)
    bool fast_norm = false;
    if constexpr (EXACT) {
        fast_norm =
            __builtin_amdgcn_readfirstlane(__syncthreads_or(unsafe ? 1 : 0) == 0 ? 1 : 0) != 0;
    } else {
        __syncthreads();
    }

#pragma unroll 1
    for (uint32_t k = 0; k < RPW; ++k) {
        if (!has_row(k)) {
            continue;
        }
        const uint32_t row = row0 + wave + k * kWaves;
        const uint32_t tw  = trial[k];
        // 2. merge + normalise + prefix sum of this lane's trial: bins 4c..4c+3
        //    of the rotated tail and head rows, one FP32 add each, then
        //    p = e * (1/sqrt(v)) (score.cpp:139-141)
        const uint32_t St = (tw >> 8) & 0xFFU;
        const uint32_t Sh = tw >> 24;
        const float2* trow = s_rows + (tw & 0xFFU) * STRIDE;
        const float2* hrow = s_rows + (a.cap + ((tw >> 16) & 0xFFU)) * STRIDE;
        float ps[NB];
        // FAST: every variance sum of the block is inside the short exact
        // sequences' range (no per-chunk range check)
        auto prefix = [&]<bool FAST>() {
            float acc = 0.0F;
#pragma unroll
            for (uint32_t c = 0; c < C4; ++c) {
                // this chunk's read offsets formed here, in each of the two
                // instances (not hoisted ahead of the branch for all chunks)
                uint32_t st = St, sh = Sh;
                __asm__ volatile("" : "+v"(st), "+v"(sh));
                const float2* tp = trow + wrap<NB>(4 * c + NB - st);
                const float2* hp = hrow + wrap<NB>(4 * c + NB - sh);
                const float2 t0 = tp[0], t1 = tp[1], t2 = tp[2], t3 = tp[3];
                const float2 h0 = hp[0], h1 = hp[1], h2 = hp[2], h3 = hp[3];
                float4 e, v;
                e.x = t0.x + h0.x;
                e.y = t1.x + h1.x;
                e.z = t2.x + h2.x;
                e.w = t3.x + h3.x;
                v.x = t0.y + h0.y;
                v.y = t1.y + h1.y;
                v.z = t2.y + h2.y;
                v.w = t3.y + h3.y;
                if constexpr (kScaledHalf<TI>) {
                    e.x *= a.unscale;
                    e.y *= a.unscale;
                    e.z *= a.unscale;
                    e.w *= a.unscale;
                    v.x *= a.unscale;
                    v.y *= a.unscale;
                    v.z *= a.unscale;
                    v.w *= a.unscale;
                }
                float4 p;
                if constexpr (FAST) {
                    p.x = e.x * rcp_rn_normal(sqrt_rn_normal(v.x));
                    p.y = e.y * rcp_rn_normal(sqrt_rn_normal(v.y));
                    p.z = e.z * rcp_rn_normal(sqrt_rn_normal(v.z));
                    p.w = e.w * rcp_rn_normal(sqrt_rn_normal(v.w));
                } else {
                    p = normalise4v<EXACT>(e, v);
                }
                acc           = (c == 0) ? p.x : acc + p.x;
                ps[4 * c]     = acc;
                acc           = acc + p.y;
                ps[4 * c + 1] = acc;
                acc           = acc + p.z;
                ps[4 * c + 2] = acc;
                acc           = acc + p.w;
                ps[4 * c + 3] = acc;
                // materialise this chunk's prefix values before the next chunk's
                // reads: left pending (the compiler sinks the products and adds),
                // every chunk's operands stay live and the kernel needs 256 VGPRs
                __asm__ volatile("" : "+v"(ps[4 * c]), "+v"(ps[4 * c + 1]), "+v"(ps[4 * c + 2]),
                                 "+v"(ps[4 * c + 3]), "+v"(acc));
            }
        };
        if (EXACT && fast_norm) {
            prefix.template operator()<true>();
        } else {
            prefix.template operator()<false>();
        }

        // 3. boxcar S/N over the widths (score.cpp:95-172), windows in registers
        const float sum = ps[NB - 1];
        const uint64_t c_idx = static_cast<uint64_t>(row) * n_f + jf;
        float best           = -FLT_MAX;
        for (uint32_t iw = 0; iw < a.wt.n; ++iw) {
            // keep the windows inside the width loop: every prefix value is opaque
            // per iteration, so no window difference of any width can be hoisted
            // (hoisted, the windows of all NB - 1 widths would be formed for every
            // trial and held in registers)
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                __asm__ volatile("" : "+v"(ps[q]));
            }
            const float dmax =
                window_max_reg<NB>(std::make_integer_sequence<int, NB - 1>{}, ps, a.wt.w[iw]);
            const float snr = (a.wt.hb[iw] * dmax) - (a.wt.b[iw] * sum);
            if constexpr (FINDMAX) {
                best = (best < snr) ? snr : best;
            } else if (a.fo.count != nullptr) {
                filter_emit(a.fo, snr > a.fo.threshold,
                            static_cast<uint32_t>(c_idx * a.wt.n + iw), snr);
            } else {
                a.scores[c_idx * a.wt.n + iw] = snr;
            }
        }
        if constexpr (FINDMAX) {
            if (a.fo.count != nullptr) {
                filter_emit(a.fo, best > a.fo.threshold, static_cast<uint32_t>(c_idx), best);
            } else {
                a.scores[c_idx] = best;
            }
        }
    }
}

/// Widest parent range (rows per side) over every block of the launch: the
/// parent index is non-decreasing in the frequency along a grid row, so a
/// block's range is spanned by the first and last frequency of its rows.
template <int NP>
__global__ void __launch_bounds__(256)
k_final2_ranges(Final2Args a, uint32_t nblocks, uint32_t* out) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nblocks) {
        return;
    }
    const uint32_t tile  = id / a.n_chunks;
    const uint32_t chunk = id % a.n_chunks;
    const uint32_t n_f   = a.lv.n_cur[2];
    const uint32_t j0    = tile * kLanes;
    const uint32_t j1    = min(n_f, j0 + kLanes) - 1;
    uint32_t lt = 0xFFFFFFFFU, ht = 0, lh = 0xFFFFFFFFU, hh = 0;
    for (uint32_t r = 0; r < a.chunks[chunk].nrows; ++r) {
        const uint32_t row = a.chunks[chunk].row0 + r;
        const FFACoord c0  = resolve_pair_rc<NP>(a.lv, a.fg.dv, row, j0);
        const FFACoord c1  = resolve_pair_rc<NP>(a.lv, a.fg.dv, row, j1);
        lt = min(lt, min(c0.i_tail, c1.i_tail));
        ht = max(ht, max(c0.i_tail, c1.i_tail));
        lh = min(lh, min(c0.i_head, c1.i_head));
        hh = max(hh, max(c0.i_head, c1.i_head));
    }
    const uint32_t m = max(ht - lt + 1, hh - lh + 1);
    atomicMax(out, m);
}

template <int NB, int NP, int RPW, typename TI>
void launch4(const Final2Args& a, uint32_t nblocks, size_t lds, bool find_max, bool exact,
             hipStream_t s) {
    if (find_max && exact) {
        k_final2<NB, NP, RPW, true, true, TI><<<nblocks, kBlock, lds, s>>>(a);
    } else if (find_max) {
        k_final2<NB, NP, RPW, true, false, TI><<<nblocks, kBlock, lds, s>>>(a);
    } else if (exact) {
        k_final2<NB, NP, RPW, false, true, TI><<<nblocks, kBlock, lds, s>>>(a);
    } else {
        k_final2<NB, NP, RPW, false, false, TI><<<nblocks, kBlock, lds, s>>>(a);
    }
}

template <int NB, int NP, typename TI>
void launch3(const Final2Args& a, uint32_t nblocks, bool find_max, bool exact, hipStream_t s) {
    const size_t lds = static_cast<size_t>(2) * a.cap * kStride<NB> * sizeof(float2);
    if (a.rows_per_block > kWaves) {
        launch4<NB, NP, kRowsPerWave, TI>(a, nblocks, lds, find_max, exact, s);
    } else {
        launch4<NB, NP, 1, TI>(a, nblocks, lds, find_max, exact, s);
    }
}

template <int NB, typename TI>
void launch2(const Final2Args& a, uint32_t nblocks, bool find_max, bool exact, hipStream_t s) {
    switch (a.lv.np) {
    case 1: return launch3<NB, 1, TI>(a, nblocks, find_max, exact, s);
    case 2: return launch3<NB, 2, TI>(a, nblocks, find_max, exact, s);
    case 3: return launch3<NB, 3, TI>(a, nblocks, find_max, exact, s);
    default: LOKI_REQUIRE(false, "final-level kernel: np must be 1, 2 or 3");
    }
}

template <typename TI>
void launch1(const Final2Args& a, uint32_t nblocks, bool find_max, bool exact, hipStream_t s) {
    switch (a.lv.nbins) {
    case 8: return launch2<8, TI>(a, nblocks, find_max, exact, s);
    case 16: return launch2<16, TI>(a, nblocks, find_max, exact, s);
    case 32: return launch2<32, TI>(a, nblocks, find_max, exact, s);
    default: LOKI_REQUIRE(false, "final-level kernel: nbins must be 8, 16 or 32");
    }
}

}  // namespace

void final2_ranges(const Final2Args& a, uint32_t nblocks, uint32_t* out, hipStream_t s) {
    LOKI_REQUIRE(out != nullptr && a.chunks != nullptr && a.n_chunks > 0, "null argument");
    if (nblocks == 0) {
        return;
    }
    const uint32_t grid = (nblocks + 255) / 256;
    switch (a.lv.np) {
    case 1: k_final2_ranges<1><<<grid, 256, 0, s>>>(a, nblocks, out); break;
    case 2: k_final2_ranges<2><<<grid, 256, 0, s>>>(a, nblocks, out); break;
    case 3: k_final2_ranges<3><<<grid, 256, 0, s>>>(a, nblocks, out); break;
    default: LOKI_REQUIRE(false, "final-level kernel: np must be 1, 2 or 3");
    }
}

bool final2_supported(uint32_t nbins) { return nbins == 8 || nbins == 16 || nbins == 32; }

uint32_t final2_lds_bytes(uint32_t nbins, uint32_t cap) {
    const uint32_t stride = nbins == 32 ? kStride<32> : (nbins == 16 ? kStride<16> : kStride<8>);
    return 2U * cap * stride * static_cast<uint32_t>(sizeof(float2));
}

void final2_merge_score(const Final2Args& a, FoldStorage si, uint32_t nblocks, bool find_max,
                        bool exact, hipStream_t s) {
    LOKI_REQUIRE(final2_supported(a.lv.nbins), "final-level kernel: unsupported nbins");
    LOKI_REQUIRE(a.wt.wmax < a.lv.nbins, "boxcar widths must be < nbins");
    LOKI_REQUIRE(a.rows_per_block >= 1 && a.rows_per_block <= kMaxRows,
                 "final-level kernel: at most 16 grid rows per block");
    LOKI_REQUIRE(a.cap <= 256, "final-level kernel: parent ranges above 256 rows");
    LOKI_REQUIRE(a.n_chunks > 0 && a.chunks != nullptr && a.cap > 0,
                 "final-level kernel: no row chunks");
    LOKI_REQUIRE(a.lv.np == 1 || a.rows != nullptr, "final-level kernel: row terms missing");
    if (a.grand) {
        LOKI_REQUIRE(a.lv.np == 1 ? a.coords1f != nullptr : a.coords1 != nullptr,
                     "final-level kernel: parent shift table missing");
    }
    LOKI_REQUIRE(final2_lds_bytes(a.lv.nbins, a.cap) <= 64U * 1024U,
                 "final-level kernel: parent ranges exceed the LDS budget");
    if (nblocks == 0) {
        return;
    }
    switch (si) {
    case FoldStorage::kF32: return launch1<float>(a, nblocks, find_max, exact, s);
    case FoldStorage::kF16: return launch1<_Float16>(a, nblocks, find_max, exact, s);
    case FoldStorage::kBF16: return launch1<__bf16>(a, nblocks, find_max, exact, s);
    case FoldStorage::kBFP16:
        LOKI_REQUIRE(a.ein != nullptr, "final-level kernel: exponent plane missing");
        return launch1<int16_t>(a, nblocks, find_max, exact, s);
    }
}

}  // namespace loki_amd::kern
