// Fused last FFA level: FP64 shift tables + merge + boxcar S/N, one pass.
//
// compute_ffa_scores (ffa.cpp:444-466) writes the final fold cube and reads
// it back for snr_boxcar_3d; here the final rows never reach HBM.  Each
// 64-lane workgroup owns 64 consecutive final coordinates:
//   1. lane i resolves coordinate c0+i in FP64 (taylor_ffa.cpp:19-271 order,
//      the same bits as the table kernels) into LDS;
//   2. the wave merges the 64 rows cooperatively: nbins/4 lanes per row, each
//      lane owning 4 phase bins of both e and v, so it forms the merged e and
//      v (kernels.cpp:71-146: one FP32 add each) and directly the normalised
//      profile p = e * (1/sqrt(v)) (score.cpp:139-141); loads of several rows
//      are in flight at once;
//   3. lane i scores row i like snr_boxcar_3d (score.cpp:95-172): sequential
//      circular prefix sum, max window difference per width, snr = (h + b) *
//      dmax - b * sum.  For nbins <= 32 the prefix sum and its wrap stay in
//      VGPRs and each width's window offsets are compile-time (one unrolled
//      window per possible width, picked by the wave-uniform width), so the
//      LDS holds only the 64 profile rows -- the shift-table records live in
//      the rows' slots until the profiles replace them -- and four waves fit
//      on each SIMD; wider rows keep the prefix sum in an LDS row.
// Every FP32 operation matches the unfused path and the CPU oracle bit for
// bit.  With FP16/BF16 parents (storage.hpp) the loads are 8-byte chunks
// widened to FP32 and the merged sums are unscaled exactly before step 2's
// normalisation; block floating-point parents are widened with their rows'
// exponents and need no unscaling.
#include <algorithm>
#include <cfloat>
#include <type_traits>
#include <utility>

#include "final_device.hpp"
#include "merge_device.hpp"
#include "resolve_device.hpp"
#include "score_device.hpp"

namespace loki_amd::kern {
namespace {

constexpr uint32_t kRows     = 64;

// Child rows scheduled back to back and frequency tiles per run: all rows of
// the last level (up to 128) sweep one narrow frequency window together, so
// the parent rows they share stay in L2 (2^23-sample accel+freq search, 77
// rows: 10.69 ms at 16 x 16, 10.2 ms at 77 x 4).
constexpr uint32_t kRowGroup = 128;
constexpr uint32_t kRun      = 4;

template <int NB, int NP, bool FINDMAX, bool EXACT, typename TI>
__device__ __forceinline__ void final_block(uint32_t bid, const TI* __restrict__ in,
                                            uint64_t nprev, const ResolveLevel& lv,
                                            const FinalGrid& fg, const WidthTable& wt,
                                            float* __restrict__ scores, const TileSchedule& ts,
                                            float unscale, const FilterOut& fo,
                                            const uint16_t* __restrict__ ein) {
    using CI = chunk4_t<TI>;
    constexpr uint32_t C4  = NB / 4;        // float4 chunks per component row
    constexpr uint32_t LPR = C4;            // lanes per row (4 bins of e and v each)
    constexpr uint32_t RPI = kRows / LPR;   // rows per merge step
    constexpr uint32_t NSTEP = kRows / RPI; // merge steps for a full tile
    constexpr uint32_t U   = NSTEP >= 4 ? 4 : NSTEP;  // row steps with loads in flight
    extern __shared__ float s_p[];          // [kRows][stride]
    // With register windows a row's LDS slots hold its shift-table record
    // until its profile replaces it (row r's record is read, in program
    // order, before any lane stores row r), so the workgroup needs no LDS
    // beyond the kRows x stride profile rows.
    constexpr bool kCoordInRow = NB <= kRegWindowMaxNB;
    __shared__ FFACoord s_coord[kCoordInRow ? 1 : kRows];
    const uint32_t stride = fg.stride;
    auto coord_store = [&](uint32_t r, const FFACoord& cd) {
        if constexpr (kCoordInRow) {
            auto* q = reinterpret_cast<uint32_t*>(s_p + r * stride);
            q[0] = cd.i_tail;
            q[1] = __float_as_uint(cd.shift_tail);
            q[2] = cd.i_head;
            q[3] = __float_as_uint(cd.shift_head);
        } else {
            s_coord[r] = cd;
        }
    };
    auto coord_load = [&](uint32_t r) {
        if constexpr (kCoordInRow) {
            const auto* q = reinterpret_cast<const uint32_t*>(s_p + r * stride);
            FFACoord cd;
            cd.i_tail     = q[0];
            cd.shift_tail = __uint_as_float(q[1]);
            cd.i_head     = q[2];
            cd.shift_head = __uint_as_float(q[3]);
            return cd;
        } else {
            return s_coord[r];
        }
    };

    const Tile tl = tile_of(ts, xcd_remap(bid, ts.nblocks), kRows);
    if (tl.count == 0) {
        return;  // idle block of a partial row group (before any barrier)
    }
    const uint32_t lane = threadIdx.x;
    const uint64_t c    = tl.c0 + lane;
    const uint32_t nrow = tl.count;

    // 1. shift tables for this tile (FP64)
    if (lane < nrow) {
        coord_store(lane, resolve_pair<NP>(lv, fg.dv, c));
    }
    __syncthreads();

    // 2. merge + normalise: lane (rsub, i4) produces p[4*i4 .. 4*i4+3] of
    //    rows rsub, rsub + RPI, ...
    {
        const uint32_t rsub = lane / LPR;
        const uint32_t i4   = lane % LPR;
#pragma unroll
        for (uint32_t s0 = 0; s0 < NSTEP; s0 += U) {
            float4 te[U], tv[U], he[U], hv[U];
            uint32_t b0[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t r = rsub + (s0 + u) * RPI;
                if (r < nrow) {
                    const FFACoord cd = coord_load(r);
                    const CI* __restrict__ trow = reinterpret_cast<const CI*>(
                        in + static_cast<uint64_t>(cd.i_tail) * (2 * NB));
                    const CI* __restrict__ hrow = reinterpret_cast<const CI*>(
                        in + (nprev + static_cast<uint64_t>(cd.i_head)) * (2 * NB));
                    // child[j] = tail[j - St] + head[j - Sh] (mod NB).  The tail is
                    // read unrotated and the sums stored at j = 4*i4 + k + St, so
                    // only the head is rotated, by Sh - St: the same two operands
                    // of every add, half the rotations.
                    const uint32_t St = NP == 1 ? 0U : shift_int(cd.shift_tail, NB);
                    const uint32_t Sh = shift_int(cd.shift_head, NB);
                    float ste = 1.0F, stv = 1.0F, she = 1.0F, shv = 1.0F;
                    if constexpr (kBlockFP<TI>) {
                        ste = bfp_row_scale(ein, cd.i_tail, 0);
                        stv = bfp_row_scale(ein, cd.i_tail, 1);
                        she = bfp_row_scale(ein, nprev + cd.i_head, 0);
                        shv = bfp_row_scale(ein, nprev + cd.i_head, 1);
                    }
                    te[u] = chunk4_st<TI>(trow, i4, ste);
                    tv[u] = chunk4_st<TI>(trow + C4, i4, stv);
                    he[u] = rotated4_st<NB, TI>(hrow, i4, (Sh - St) & (NB - 1), she);
                    hv[u] = rotated4_st<NB, TI>(hrow + C4, i4, (Sh - St) & (NB - 1), shv);
                    b0[u] = 4 * i4 + St;
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t r = rsub + (s0 + u) * RPI;
                if (r < nrow) {
                    merge_store<NB, EXACT, TI>(te[u], tv[u], he[u], hv[u], unscale,
                                               s_p + r * stride, b0[u]);
                }
            }
        }
    }
    __syncthreads();
    if (lane >= nrow) {
        return;
    }

    // 3. boxcar S/N of row `lane`: circular prefix sum ps and its wrap
    //    ext[k] = ps[k] + sum (score.cpp:143-156).  Up to kRegWindowMaxNB bins
    //    both stay in registers; wider rows become the row's LDS slots
    //    (n + wmax entries, stride >= n + wmax).
    float* __restrict__ row = s_p + lane * stride;
    float ps[NB];
    float acc = 0.0F;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const float x = row[j];
        acc           = (j == 0) ? x : acc + x;
        ps[j]         = acc;
    }
    const float sum = ps[NB - 1];
    [[maybe_unused]] float ext[NB - 1];
    if constexpr (NB <= kRegWindowMaxNB) {
#pragma unroll
        for (int k = 0; k < NB - 1; ++k) {
            ext[k] = ps[k] + sum;
        }
    } else {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            row[j] = ps[j];
        }
        for (uint32_t j = 0; j < wt.wmax; ++j) {
            row[NB + j] = row[j] + sum;
        }
    }
    float best = -FLT_MAX;
    for (uint32_t iw = 0; iw < wt.n; ++iw) {
        float dmax;
        if constexpr (NB <= kRegWindowMaxNB) {
            // The windows do not depend on the width loop; hoisted out of it,
            // all NB - 1 of them would be formed for every row.  An empty
            // volatile asm on one operand ties them to this iteration.
            __asm__ volatile("" : "+v"(ps[0]));
            dmax = window_max_reg<NB>(std::make_integer_sequence<int, NB - 1>{}, ps, ext,
                                      wt.w[iw]);
        } else {
            dmax = window_max<NB>(row + wt.w[iw], ps);
        }
        const float snr  = ((wt.h[iw] + wt.b[iw]) * dmax) - (wt.b[iw] * sum);
        if constexpr (FINDMAX) {
            best = (best < snr) ? snr : best;
        } else if (fo.count != nullptr) {
            filter_emit(fo, snr > fo.threshold, static_cast<uint32_t>(c * wt.n + iw), snr);
        } else {
            scores[c * wt.n + iw] = snr;
        }
    }
    if constexpr (FINDMAX) {
        if (fo.count != nullptr) {
            filter_emit(fo, best > fo.threshold, static_cast<uint32_t>(c), best);
        } else {
            scores[c] = best;
        }
    }
}

template <int NB, int NP, bool FINDMAX, bool EXACT, typename TI>
__global__ void __launch_bounds__(kRows)
k_final_merge_score(const TI* __restrict__ in, uint64_t nprev, ResolveLevel lv, FinalGrid fg,
                    WidthTable wt, float* __restrict__ scores, TileSchedule ts, float unscale,
                    FilterOut fo, const uint16_t* __restrict__ ein) {
    final_block<NB, NP, FINDMAX, EXACT, TI>(blockIdx.x, in, nprev, lv, fg, wt, scores, ts,
                                            unscale, fo, ein);
}

// Grouped launch: the last levels of many independent FFAs (one task each,
// sorted by first block); block -> task by binary search over block0.
template <int NB, int NP, bool FINDMAX, bool EXACT, typename TI>
__global__ void __launch_bounds__(kRows)
k_final_grouped(const FinalTask* __restrict__ tasks, uint32_t ntasks) {
    static_assert(kRows == kFinalRows, "grouped tiles must match the single-level kernel");
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
    const FinalTask& t = tasks[lo];
    final_block<NB, NP, FINDMAX, EXACT, TI>(blockIdx.x - t.block0, static_cast<const TI*>(t.in),
                                            t.nprev, t.lv, t.fg, t.wt, t.scores, t.ts,
                                            t.unscale, FilterOut{}, nullptr);
}

template <int NB, int NP, typename TI>
void launch_nb_np(const TI* in, float unscale, uint64_t ncur, uint64_t nprev,
                  const ResolveLevel& lv, const WidthTable& wt, float* scores, bool find_max,
                  bool exact, const FilterOut& fo, hipStream_t s) {
    // nseg == 1 at the last level; child rows sharing parents run back to back
    const TileSchedule ts = make_schedule(ncur, lv.n_cur[2], 1, kRows, kRowGroup, kRun);
    const uint32_t nb     = ts.nblocks;
    FinalGrid fg{};
    for (int k = 0; k < 3; ++k) {
        fg.dv[k] = (lv.hi[k] - lv.lo[k]) / static_cast<double>(lv.n_cur[k]);
    }
    fg.stride = final_lds_stride(NB, wt.wmax);
    const size_t lds = static_cast<size_t>(kRows) * fg.stride * sizeof(float);
    // BFP16 parents: the exponent plane follows the level's 2 * nprev rows
    const uint16_t* ein = nullptr;
    if constexpr (kBlockFP<TI>) {
        ein = reinterpret_cast<const uint16_t*>(in + 2 * nprev * (2 * NB));
    }
    if (find_max && exact) {
        k_final_merge_score<NB, NP, true, true, TI>
            <<<nb, kRows, lds, s>>>(in, nprev, lv, fg, wt, scores, ts, unscale, fo, ein);
    } else if (find_max) {
        k_final_merge_score<NB, NP, true, false, TI>
            <<<nb, kRows, lds, s>>>(in, nprev, lv, fg, wt, scores, ts, unscale, fo, ein);
    } else if (exact) {
        k_final_merge_score<NB, NP, false, true, TI>
            <<<nb, kRows, lds, s>>>(in, nprev, lv, fg, wt, scores, ts, unscale, fo, ein);
    } else {
        k_final_merge_score<NB, NP, false, false, TI>
            <<<nb, kRows, lds, s>>>(in, nprev, lv, fg, wt, scores, ts, unscale, fo, ein);
    }
}

template <int NB, typename TI>
void launch_nb(const TI* in, float unscale, uint64_t ncur, uint64_t nprev,
               const ResolveLevel& lv, const WidthTable& wt, float* scores, bool find_max,
               bool exact, const FilterOut& fo, hipStream_t s) {
    switch (lv.np) {
    case 1:
        return launch_nb_np<NB, 1>(in, unscale, ncur, nprev, lv, wt, scores, find_max, exact, fo,
                                   s);
    case 2:
        return launch_nb_np<NB, 2>(in, unscale, ncur, nprev, lv, wt, scores, find_max, exact, fo,
                                   s);
    case 3:
        return launch_nb_np<NB, 3>(in, unscale, ncur, nprev, lv, wt, scores, find_max, exact, fo,
                                   s);
    default: LOKI_REQUIRE(false, "final-level kernel: np must be 1, 2 or 3");
    }
}

template <typename TI>
void launch_typed(const void* in, float unscale, uint64_t ncur, uint64_t nprev,
                  const ResolveLevel& lv, const WidthTable& wt, float* scores, bool find_max,
                  bool exact, const FilterOut& fo, hipStream_t s) {
    const auto* p = static_cast<const TI*>(in);
    switch (lv.nbins) {
    case 8: return launch_nb<8>(p, unscale, ncur, nprev, lv, wt, scores, find_max, exact, fo, s);
    case 16:
        return launch_nb<16>(p, unscale, ncur, nprev, lv, wt, scores, find_max, exact, fo, s);
    case 32:
        return launch_nb<32>(p, unscale, ncur, nprev, lv, wt, scores, find_max, exact, fo, s);
    case 64:
        return launch_nb<64>(p, unscale, ncur, nprev, lv, wt, scores, find_max, exact, fo, s);
    case 128:
        return launch_nb<128>(p, unscale, ncur, nprev, lv, wt, scores, find_max, exact, fo, s);
    default: break;
    }
}

}  // namespace

bool final_merge_score_supported(uint32_t nbins) {
    return nbins == 8 || nbins == 16 || nbins == 32 || nbins == 64 || nbins == 128;
}

void final_merge_score(const void* in, FoldStorage si, float unscale, uint64_t ncur,
                       uint64_t nprev, const ResolveLevel& lv, const WidthTable& wt,
                       float* scores, bool find_max, bool exact, hipStream_t s,
                       const FilterOut* fo) {
    LOKI_REQUIRE(final_merge_score_supported(lv.nbins), "final-level kernel: unsupported nbins");
    LOKI_REQUIRE(wt.wmax < lv.nbins, "boxcar widths must be < nbins");
    if (ncur == 0) {
        return;
    }
    const FilterOut f = fo != nullptr ? *fo : FilterOut{};
    if (fo != nullptr) {
        LOKI_REQUIRE(fo->count != nullptr, "score filter: null counter");
        LOKI_REQUIRE(ncur * (find_max ? 1U : wt.n) <= 0xFFFFFFFFULL,
                     "score filter: trial index exceeds 32 bits");
    }
    switch (si) {
    case FoldStorage::kF32:
        return launch_typed<float>(in, unscale, ncur, nprev, lv, wt, scores, find_max, exact, f,
                                   s);
    case FoldStorage::kF16:
        return launch_typed<_Float16>(in, unscale, ncur, nprev, lv, wt, scores, find_max, exact,
                                      f, s);
    case FoldStorage::kBF16:
        return launch_typed<__bf16>(in, unscale, ncur, nprev, lv, wt, scores, find_max, exact, f,
                                    s);
    case FoldStorage::kBFP16:
        return launch_typed<int16_t>(in, unscale, ncur, nprev, lv, wt, scores, find_max, exact,
                                     f, s);
    }
}

namespace {

template <int NB, typename TI>
void grouped_launch(const FinalTask* tasks, uint32_t ntasks, uint32_t nblocks, bool find_max,
                    bool exact, size_t lds, hipStream_t s) {
    if (find_max && exact) {
        k_final_grouped<NB, 1, true, true, TI><<<nblocks, kRows, lds, s>>>(tasks, ntasks);
    } else if (find_max) {
        k_final_grouped<NB, 1, true, false, TI><<<nblocks, kRows, lds, s>>>(tasks, ntasks);
    } else if (exact) {
        k_final_grouped<NB, 1, false, true, TI><<<nblocks, kRows, lds, s>>>(tasks, ntasks);
    } else {
        k_final_grouped<NB, 1, false, false, TI><<<nblocks, kRows, lds, s>>>(tasks, ntasks);
    }
}

}  // namespace

namespace {

// Grouped launch of final_tile: 256-lane workgroups, WS lanes per trial; the
// tiles of each task in XCD-remapped order (as final_block's).
template <int NB, int WS>
__global__ void __launch_bounds__(kFinalWideBlock)
k_final_grouped_wide(const FinalTask* __restrict__ tasks, uint32_t ntasks) {
    extern __shared__ __align__(16) unsigned char s_wide[];
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
    const FinalTask& t = tasks[lo];
    final_tile<NB, WS>(xcd_remap(blockIdx.x - t.block0, t.ts.nblocks), t, s_wide);
}

template <int NB>
void wide_launch(const FinalTask* tasks, uint32_t ntasks, uint32_t nblocks, uint32_t lanes,
                 size_t lds, hipStream_t s) {
    if constexpr (NB >= 16) {
        if (lanes == 4) {
            k_final_grouped_wide<NB, 4><<<nblocks, kFinalWideBlock, lds, s>>>(tasks, ntasks);
            return;
        }
    }
    if constexpr (NB <= 64) {
        if (lanes == 2) {
            k_final_grouped_wide<NB, 2><<<nblocks, kFinalWideBlock, lds, s>>>(tasks, ntasks);
            return;
        }
    }
    if constexpr (NB <= 32) {
        if (lanes == 1) {
            k_final_grouped_wide<NB, 1><<<nblocks, kFinalWideBlock, lds, s>>>(tasks, ntasks);
            return;
        }
    }
    LOKI_REQUIRE(false, "grouped last level: unsupported lanes per trial for this nbins");
}

}  // namespace

uint32_t final_wide_rows(uint32_t lanes) { return kFinalWideBlock / std::max(lanes, 1U); }

bool final_wide_supported(uint32_t nbins, uint32_t lanes) {
    const bool nb_ok = nbins == 8 || nbins == 16 || nbins == 32 || nbins == 64 || nbins == 128;
    const bool ws_ok = lanes == 1 || lanes == 2 || lanes == 4;
    // whole merge steps per tile (256 / lanes rows >= 1024 / nbins rows per
    // step) and at most 64 KiB of profile rows (wider rows need more lanes)
    return nb_ok && ws_ok && nbins >= 4 * lanes && (nbins <= 32 || nbins <= 32 * lanes);
}

size_t final_wide_lds_bytes(uint32_t lanes, uint32_t stride) {
    const size_t rows = final_wide_rows(lanes);
    return rows * stride * sizeof(float) + rows * sizeof(FFACoord);
}

void final_grouped_wide(const FinalTask* tasks, uint32_t ntasks, uint32_t nblocks,
                        uint32_t nbins, uint32_t lanes, size_t lds, hipStream_t s) {
    LOKI_REQUIRE(final_wide_supported(nbins, lanes), "grouped last level: unsupported lanes");
    LOKI_REQUIRE(lds <= 64 * 1024, "grouped last level: more than 64 KiB of LDS");
    if (ntasks == 0 || nblocks == 0) {
        return;
    }
    switch (nbins) {
    case 8: return wide_launch<8>(tasks, ntasks, nblocks, lanes, lds, s);
    case 16: return wide_launch<16>(tasks, ntasks, nblocks, lanes, lds, s);
    case 32: return wide_launch<32>(tasks, ntasks, nblocks, lanes, lds, s);
    case 64: return wide_launch<64>(tasks, ntasks, nblocks, lanes, lds, s);
    case 128: return wide_launch<128>(tasks, ntasks, nblocks, lanes, lds, s);
    default: LOKI_REQUIRE(false, "grouped last level: nbins must be 8, 16, 32, 64 or 128");
    }
}

void final_grouped(const FinalTask* tasks, uint32_t ntasks, uint32_t nblocks, uint32_t nbins,
                   FoldStorage si, bool find_max, bool exact, uint32_t lds_stride,
                   hipStream_t s) {
    LOKI_REQUIRE(si == FoldStorage::kF32, "grouped last level: FP32 parents only");
    if (ntasks == 0 || nblocks == 0) {
        return;
    }
    const size_t lds = static_cast<size_t>(kRows) * lds_stride * sizeof(float);
    switch (nbins) {
    case 8: return grouped_launch<8, float>(tasks, ntasks, nblocks, find_max, exact, lds, s);
    case 16: return grouped_launch<16, float>(tasks, ntasks, nblocks, find_max, exact, lds, s);
    case 32: return grouped_launch<32, float>(tasks, ntasks, nblocks, find_max, exact, lds, s);
    case 64: return grouped_launch<64, float>(tasks, ntasks, nblocks, find_max, exact, lds, s);
    case 128: return grouped_launch<128, float>(tasks, ntasks, nblocks, find_max, exact, lds, s);
    default: LOKI_REQUIRE(false, "grouped last level: nbins must be 8, 16, 32, 64 or 128");
    }
}

uint32_t final_lds_stride(uint32_t nbins, uint32_t wmax) {
    // odd: conflict-free rows; n profile slots (windows in registers) or
    // n + wmax prefix-sum slots
    return (nbins <= static_cast<uint32_t>(kRegWindowMaxNB) ? nbins : nbins + wmax) | 1U;
}

}  // namespace loki_amd::kern
