// Host-side launchers of the gfx950 HIP kernels.  Every launcher is
// asynchronous on the given stream and validates shapes before launching.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "device.hpp"
#include "merge_device.hpp"
#include "plan.hpp"
#include "storage.hpp"

namespace loki_amd::kern {

/// Geometry of one merge level for the shift-table resolvers.  The last
/// three search parameters are (jerk, accel, freq); unused leading entries
/// have count 1.
struct ResolveLevel {
    uint32_t n_cur[3];
    uint32_t n_prev[3];
    double lo[3];
    double hi[3];
    double tseg_prev;  ///< tseg_brute * 2^(level-1)
    uint32_t nbins;
    uint32_t np;       ///< 1 (freq), 2 (accel, freq) or 3 (jerk, accel, freq)
};

/// Row part of the last level's resolver for one side of one grid row
/// (flattened (jerk, accel) cell): f_new = f_cur * fac, the phase total
/// td * f_cur (td = delta_t - delay_rel) and the parent index base + idx_f.
struct RowTerms {
    double fac;
    double td;
    uint32_t base;
    uint32_t pad;
};

/// Extra per-level constants of the fused last-level kernel.
struct FinalGrid {
    double dv[3];     ///< (hi - lo) / n_cur per parameter (psr_utils.hpp:290-291)
    uint32_t stride;  ///< LDS row stride in floats (>= nbins, odd)
};

/// Fills ncoords FFACoordFreq records (np == 1) of one level.
void resolve_freq(const ResolveLevel& lv, FFACoordFreq* out, uint64_t ncoords, hipStream_t s);
/// Fills ncoords FFACoord records (np == 2 or 3) of one level, tail and head,
/// from the level's row terms `rows` (device; tail and head per grid row).
void resolve_poly(const ResolveLevel& lv, const RowTerms* rows, FFACoord* out, uint64_t ncoords,
                  hipStream_t s);

/// Bucket lists of the time-domain brute fold, built once per plan: for
/// each frequency f and phase bin b, the sample indices of the bin in
/// ascending order, stored transposed: list[base[f] + k * nbins + b].
/// Segments longer than the LDS staging limit use no lists: their buckets are
/// walked as contiguous runs of samples (use_runs).
/// Bucket lists of a brute fold as device pointers (owned elsewhere).
struct BruteFoldView {
    uint32_t nfreqs{0};
    uint32_t seglen{0};
    uint32_t nbins{0};
    bool use_runs{false};
    double tsamp{0.0};
    double t_ref{0.0};
    const double* freqs{nullptr};    ///< f64 [nfreqs]
    const uint32_t* counts{nullptr};  ///< u32 [nfreqs * nbins]
    const uint64_t* base{nullptr};    ///< u64 [nfreqs]
    const uint32_t* list{nullptr};    ///< u32 [sum_f kmax_f * nbins]
};
struct BruteFoldLists {
    uint32_t nfreqs{0};
    uint32_t seglen{0};
    uint32_t nbins{0};
    bool use_runs{false};
    double tsamp{0.0};
    double t_ref{0.0};
    DeviceBuffer freqs;   ///< f64 [nfreqs]
    DeviceBuffer counts;  ///< u32 [nfreqs * nbins]
    DeviceBuffer base;    ///< u64 [nfreqs]
    DeviceBuffer list;    ///< u32 [sum_f kmax_f * nbins]
    BruteFoldView view() const {
        return {nfreqs, seglen, nbins, use_runs, tsamp, t_ref,
                freqs.as<double>(), counts.as<uint32_t>(), base.as<uint64_t>(),
                list.as<uint32_t>()};
    }
};
/// FP64 phase bins (psr_utils.hpp:53-63) -> lists; synchronises `s`.
BruteFoldLists build_brute_fold_lists(const double* freqs_dev, uint32_t nfreqs, uint32_t seglen,
                                      uint32_t nbins, double tsamp, double t_ref, hipStream_t s);
/// fold[nseg, nfreqs, 2, nbins] = ascending-order bucket sums of each segment
/// (FP32), stored as `so`; unless so is FP32 each sum is multiplied by
/// `scale` before rounding.
void brute_fold_time(const float* ts_e, const float* ts_v, uint64_t nseg,
                     const BruteFoldLists& lists, void* fold, FoldStorage so, float scale,
                     hipStream_t s);
void brute_fold_time(const float* ts_e, const float* ts_v, uint64_t nseg,
                     const BruteFoldView& lists, void* fold, FoldStorage so, float scale,
                     hipStream_t s);
/// Bytes of the brute fold's file-scope device scratch (allocation-free entry point).
uint64_t brute_fold_scratch_bytes();
/// FP32 time brute fold of nsamps / seglen segments with its frequencies,
/// bucket counts, offsets and lists in the file-scope scratch array instead
/// of allocated buffers: allocates no device memory.  Throws if they need
/// more than brute_fold_scratch_bytes().  Synchronises `s`.  Calls that
/// overlap on one device race on the scratch.
void brute_fold_time_noalloc(const float* ts_e, const float* ts_v, uint64_t nsamps,
                             const double* freqs_host, uint32_t nfreqs, uint32_t seglen,
                             uint32_t nbins, double tsamp, double t_ref, float* fold,
                             hipStream_t s);

/// Time-domain merge of one level (freq-only when coords_freq != nullptr);
/// n_f: frequency grid points of the level (coordinates are rows of n_f).
/// si/so: storage of the parent/child level (so == si, or FP32 to return a
/// reduced-precision tree to full precision); unless both are FP32 the FP32
/// sum is multiplied by `scale` before rounding to so.
void merge_time(const void* in, void* out, const FFACoord* coords,
                const FFACoordFreq* coords_freq, uint64_t ncur, uint64_t nprev, uint64_t nseg,
                uint64_t n_f, uint32_t nbins, FoldStorage si, FoldStorage so, float scale,
                hipStream_t s);

/// Fourier brute fold: fold [nseg, nfreqs, 2, nbins/2+1] complex (float2) of
/// exp(-2 pi i m f (k tsamp - t_ref)) sums; FP32 matrix-core DFT with FP64
/// phases, harmonic 0 by FP64 reduction.  nbins in [2, 256].
/// split_bf16: harmonics on bf16 matrix cores with each operand split into
/// high and low bf16 parts (three products per term, FP32 accumulation) for
/// nbins <= 129; otherwise FP32 matrix cores.
void brute_fold_fourier(const float* ts_e, const float* ts_v, uint64_t nseg, uint32_t seglen,
                        const double* freqs_dev, uint32_t nfreqs, uint32_t nbins, double tsamp,
                        double t_ref, float2* fold, hipStream_t s, bool split_bf16 = false);

/// Fourier-domain merge of one level (freq-only when coords_freq != nullptr):
/// phase rotation by the fractional shifts, rows of [2, nbins/2+1] float2.
void merge_fourier(const float2* in, float2* out, const FFACoord* coords,
                   const FFACoordFreq* coords_freq, uint64_t ncur, uint64_t nprev, uint64_t nseg,
                   uint32_t nbins, hipStream_t s);

/// Inverse real DFT of nrows rows of nbins/2+1 float2 into rows of nbins
/// floats, scaled by 1/nbins (C2R + normalisation).
void irfft_rows(const float2* in, float* out, uint64_t nrows, uint32_t nbins, hipStream_t s);
/// Forward real DFT (unnormalised) of nrows rows of nbins floats.
void rfft_rows(const float* in, float2* out, uint64_t nrows, uint32_t nbins, hipStream_t s);

/// Workspace bytes of refine_forward / refine_adjoint (either may reuse it).
SizeType refine_workspace_bytes(uint64_t nsamps, uint64_t ncand, uint32_t nbins);
/// Continuous Fourier-fold score of ncand candidates theta [ncand][3] =
/// (f, fdot, fddot) (FP64, device) over the whole series: scores [ncand] and,
/// when state != nullptr, the adjoint state [ncand][2][nbins/2+1] float2.
/// nbins in {16, 32, 64, 128}, width in [1, nbins).
void refine_forward(const float* ts_e, const float* ts_v, uint64_t nsamps, double tsamp,
                    double t_ref, const double* theta, uint64_t ncand, uint32_t nbins,
                    uint32_t width, float* scores, float2* state, void* workspace,
                    hipStream_t s);
/// dS/dtheta [ncand][3] (FP64) from the state of refine_forward at the same theta.
void refine_adjoint(const float* ts_e, const float* ts_v, uint64_t nsamps, double tsamp,
                    double t_ref, const double* theta, uint64_t ncand, uint32_t nbins,
                    const float2* state, double* grad, void* workspace, hipStream_t s);

/// Compacted output of the score-and-filter modes: (index, snr) of every
/// (trial, width) with snr > threshold (index = trial * nwidths + width), or
/// of every trial whose best S/N passes (index = trial).  `count` (device,
/// zeroed before the launch) receives the number that passed, which may exceed
/// `capacity`; only the first `capacity` slots are written, in no fixed order.
struct FilterOut {
    float threshold;
    uint32_t capacity;
    uint32_t* count;
    uint32_t* indices;
    float* snr;
};
/// Zeroes a device counter with a kernel (capturable in a graph).
void zero_counter(uint32_t* count, hipStream_t s);

/// Checks the short correctly rounded sequences of exact scoring
/// (sqrt_rn_normal, rcp_rn_normal) against the IEEE sqrtf and 1.0f / x for
/// every float bit pattern u in [bits_lo, bits_hi) inside each sequence's
/// range.  counts[0] / counts[1] (u64, device, zeroed by the caller) receive
/// the sqrt / reciprocal mismatches and counts[2] / counts[3] the number of
/// patterns checked; first_bad[0] / [1] (u32, device, preset to 0xFFFFFFFF)
/// the lowest mismatching pattern.
void selftest_exact_norm_sweep(uint32_t bits_lo, uint32_t bits_hi, unsigned long long* counts,
                               uint32_t* first_bad, hipStream_t s);
/// For n bit patterns: out[4i..4i+3] = sqrtf(x), 1.0f / x, sqrt_rn_normal(x),
/// rcp_rn_normal(x) as computed on the device.
void selftest_exact_norm_values(const uint32_t* bits, uint64_t n, float* out, hipStream_t s);

/// Boxcar widths with the reference's template constants (score.cpp:107-116).
struct WidthTable {
    static constexpr uint32_t kMax = 64;
    uint32_t n;
    uint32_t wmax;
    uint32_t w[kMax];
    float h[kMax];
    float b[kMax];
    float hb[kMax];  ///< h + b, the same FP32 sum the kernels would form
};
WidthTable make_width_table(const SizeType* widths, SizeType nw, SizeType nbins);

/// Fused last level (nsegments == 1): FP64 shift tables + merge + boxcar S/N
/// from the parent level `in` [2, nprev, 2, nbins] straight to scores.
bool final_merge_score_supported(uint32_t nbins);
/// exact: p = e * (1/sqrtf(v)) with correctly rounded operations (bit-identical to the CPU);
/// otherwise p = e * rsq(v) with the hardware reciprocal square root.
/// si: storage of the parent level; unless FP32 the merged e and v are
/// multiplied by `unscale` (exact power of two) before normalisation.
/// With `fo`, the scores go to the compacted candidate list instead.
void final_merge_score(const void* in, FoldStorage si, float unscale, uint64_t ncur,
                       uint64_t nprev, const ResolveLevel& lv, const WidthTable& wt,
                       float* scores, bool find_max, bool exact, hipStream_t s,
                       const FilterOut* fo = nullptr);

/// Chunk of final grid rows (flattened (jerk, accel) cells) whose trials share
/// their parent rows: rows [row0, row0 + nrows), nrows <= rows_per_block.
struct RowChunk {
    uint32_t row0;
    uint32_t nrows;
};
/// Arguments of the one-lane-per-trial fused last level (k_final2.hip).
/// Parent level F-1: [2, n_p1, 2, nbins] stored as `si`.  Without `grand`,
/// `in` is that level; with it, `in` is the grandparent level F-2
/// [4, n_p2, 2, nbins] and the parent rows are merged on chip through the
/// parent level's shift table (coords1 / coords1f), scaled by scale1 and
/// rounded to the storage type like a stored level.
struct Final2Args {
    const void* in;
    const FFACoord* coords1;
    const FFACoordFreq* coords1f;
    uint64_t n_p1;
    uint64_t n_p2;
    ResolveLevel lv;
    FinalGrid fg;
    WidthTable wt;
    float* scores;
    FilterOut fo;
    const RowChunk* chunks;  ///< device
    const RowTerms* rows;    ///< device, np >= 2: (tail, head) terms of every grid row
    uint32_t n_chunks;
    uint32_t rows_per_block;  ///< most rows in a chunk (<= 16; 4 waves, 4 rows each)
    uint32_t cap;             ///< staged parent rows per side
    uint32_t grand;
    float unscale;            ///< reduced precision: final sums x 2^(k0 + F - 1)
    float scale1;             ///< reduced precision: parent merge scale (0.5)
    const uint16_t* ein;      ///< block floating point: exponent plane of `in`
    uint32_t pbfp;            ///< FP32 `in` whose on-chip parent level is block floating point
};
/// nbins supported by final2_merge_score (register windows).
bool final2_supported(uint32_t nbins);
/// LDS bytes of one block staging `cap` rows per side.
uint32_t final2_lds_bytes(uint32_t nbins, uint32_t cap);
/// out (device u32, zeroed by the caller) <- the widest parent range (rows
/// per side) of any of the nblocks blocks: the `cap` final2_merge_score needs.
void final2_ranges(const Final2Args& a, uint32_t nblocks, uint32_t* out, hipStream_t s);
/// Launch over nblocks = n_chunks * frequency tiles of 64 (tile-major).
void final2_merge_score(const Final2Args& a, FoldStorage si, uint32_t nblocks, bool find_max,
                        bool exact, hipStream_t s);

// ---------------------------------------------------------------------------
// Grouped launches: many independent FFAs (a batch of series) advanced by one
// launch per stage.  Each task owns a contiguous range of blocks starting at
// block0; tasks are sorted by block0 and a block finds its task by binary
// search.
// ---------------------------------------------------------------------------
struct ResolveTask {
    ResolveLevel lv;
    void* out;        ///< FFACoordFreq (np == 1) or FFACoord records
    uint64_t n;       ///< coordinates of the level
    uint32_t block0;  ///< first 256-lane block of the task
};
struct FoldTask {
    const float* ts_e;
    const float* ts_v;
    const uint32_t* list;    ///< bucket lists (list path)
    const uint64_t* base;
    const uint32_t* counts;
    const double* freqs;     ///< frequencies (run path)
    double tsamp, t_ref;
    void* fold;
    float scale;
    uint32_t seglen, nfreqs, nbins, bps;  ///< bps: blocks per segment
    uint32_t block0;
};
struct MergeTask {
    const void* in;
    void* out;
    const FFACoord* coords;       ///< poly levels (or nullptr)
    const FFACoordFreq* coords_freq;  ///< freq-only levels (or nullptr)
    uint64_t ncur, nprev;
    TileSchedule ts;
    float scale;
    uint32_t block0;
};
struct FinalTask {
    const void* in;
    float* scores;
    uint64_t nprev;
    ResolveLevel lv;
    FinalGrid fg;
    TileSchedule ts;
    WidthTable wt;
    float unscale;
    uint32_t block0;
};

/// Grouped FP64 shift tables (np == 1 or poly, one kind per launch).
void resolve_grouped(const ResolveTask* tasks, uint32_t ntasks, uint32_t nblocks, uint32_t np,
                     hipStream_t s);
/// Blocks per segment of the time brute fold of `lists` (list or run path;
/// tasks of one grouped launch must share the path).
uint32_t brute_fold_blocks_per_segment(const BruteFoldLists& lists);
uint32_t brute_fold_blocks_per_segment(const BruteFoldView& lists);
/// Grouped time brute fold (all tasks stored as `so`): the list path, or with
/// `runs` the long-segment run path (max_nbins: the largest nbins of the
/// tasks, which sizes its LDS).
void brute_fold_grouped(const FoldTask* tasks, uint32_t ntasks, uint32_t nblocks, bool runs,
                        uint32_t max_nbins, FoldStorage so, hipStream_t s);
/// Grouped merges of one storage type, nbins and coordinate kind.
void merge_grouped(const MergeTask* tasks, uint32_t ntasks, uint32_t nblocks, uint32_t nbins,
                   bool poly, FoldStorage st, hipStream_t s);
/// Tile geometry of merge_grouped for a given nbins: coordinates per block.
uint32_t merge_tile(uint32_t nbins);
/// nbins handled by merge_time2.
bool merge2_supported(uint32_t nbins);
/// Two time-domain merge levels in one pass: level C (n_c coordinates per
/// segment, nseg_c segments) from level A (n_a coordinates, 4 * nseg_c
/// segments) through level B (n_b coordinates) -- the bits of two merge_time
/// calls, without storing B.  coords_b/_c: B's and C's shift tables (poly, or
/// the freq-only tables); n_f: frequencies per row of C.  si: A's storage,
/// so: B's and C's (equal, or FP32 -> block floating point); scale: 0.5 per
/// level for FP16/BF16.
void merge_time2(const void* in, void* out, const FFACoord* coords_b,
                 const FFACoordFreq* coords_freq_b, const FFACoord* coords_c,
                 const FFACoordFreq* coords_freq_c, uint64_t n_c, uint64_t n_b, uint64_t n_a,
                 uint64_t nseg_c, uint64_t n_f, uint32_t nbins, FoldStorage si, FoldStorage so,
                 float scale, hipStream_t s);
/// Grouped fused last levels (freq-only) of one nbins and parent storage.
void final_grouped(const FinalTask* tasks, uint32_t ntasks, uint32_t nblocks, uint32_t nbins,
                   FoldStorage si, bool find_max, bool exact, uint32_t lds_stride,
                   hipStream_t s);
/// Coordinates per block of the fused last level, and its LDS row stride.
constexpr uint32_t kFinalRows = 64;
uint32_t final_lds_stride(uint32_t nbins, uint32_t wmax);
/// Grouped fused last levels with 256-lane workgroups and `lanes` lanes per
/// trial (1, 2 or 4), each lane taking every lanes-th width: the boxcar-width
/// schedule the batch autotuner picks per nbins.  Tiles hold
/// final_wide_rows(lanes) trials (the tasks' TileSchedules are built with that
/// tile); FP32 parents, exact normalisation, one score per width; the bits of
/// final_grouped.  lds: final_wide_lds_bytes of the launch's widest stride.
void final_grouped_wide(const FinalTask* tasks, uint32_t ntasks, uint32_t nblocks,
                        uint32_t nbins, uint32_t lanes, size_t lds, hipStream_t s);
uint32_t final_wide_rows(uint32_t lanes);
/// (nbins, lanes) built: lanes 1, 2 or 4, nbins >= 4 * lanes, and above 32
/// bins nbins <= 32 * lanes (at most 64 KiB of profile rows per workgroup).
bool final_wide_supported(uint32_t nbins, uint32_t lanes);
size_t final_wide_lds_bytes(uint32_t lanes, uint32_t stride);

// ---------------------------------------------------------------------------
// Persistent batch kernel (k_tree.hip): one launch runs every stage of every
// series of a batch.  The work is a list of items (a brute fold, a level's
// shift table, a merge level or a fused last level of one series), each a
// range of tickets, one ticket per 256-lane tile.  Resident workgroups take
// tickets in order from a counter; a tile first waits until every tile of the
// items it depends on is done (per-item counters, agent-scope acquire), and
// signals its own item when its results are written (agent-scope release).
// Items are ordered so that every dependency's tickets precede the item's:
// the smallest unfinished ticket never waits on an unfinished tile, so the
// launch always drains.
// ---------------------------------------------------------------------------
enum class TreeKind : uint32_t { kResolve = 0, kFoldList = 1, kFoldUnits = 2, kMerge = 3, kFinal = 4 };
constexpr uint32_t kTreeBlock  = 256;
constexpr uint32_t kTreeNoDep  = 0xFFFFFFFFU;
constexpr uint32_t kTreeSyncHeader = 3;  ///< ticket counter, finished workgroups, error flags
struct TreeItem {
    uint32_t kind;     ///< TreeKind
    uint32_t nbins;    ///< merge / final: bins of the series
    uint32_t task;     ///< index into the kind's task table
    uint32_t ticket0;  ///< first ticket (items sorted by ticket0)
    uint32_t ntiles;   ///< tickets of the item
    uint32_t dep[2];   ///< items whose tiles must all be done first (kTreeNoDep: none)
    uint32_t lanes;    ///< final: threads per trial in the S/N stage (1, 2, 4 or 8)
};
struct TreeArgs {
    const TreeItem* items;  ///< device
    uint32_t nitems;
    uint32_t ntickets;
    const ResolveTask* resolve;  ///< device task tables (block0 unused)
    const FoldTask* fold;
    const MergeTask* merge;
    const FinalTask* final_;
    /// device, kTreeSyncHeader + nitems words, zero before the first launch;
    /// the launch leaves [0, 2) and the item counters zero again, and ORs 1
    /// into [2] if a wait exceeded spin_limit polls (then results are invalid)
    uint32_t* sync;
    uint32_t spin_limit;
};
/// Dynamic LDS bytes of one item kind (fold: its nbins; final: nbins, lanes
/// and the LDS row stride of final_lds_stride).
size_t tree_lds_bytes(TreeKind kind, uint32_t nbins, uint32_t lanes, uint32_t stride);
/// Workgroups of 256 lanes resident at once with `lds` bytes each.
uint32_t tree_resident_blocks(size_t lds);
/// One persistent launch of nblocks workgroups over every ticket.
void tree_launch(const TreeArgs& a, uint32_t nblocks, size_t lds, hipStream_t s);

/// snr_boxcar_3d[_max]: folds [nprof, 2, nbins] -> scores [nprof, nw] or [nprof].
/// With `fo`, the scores go to the compacted candidate list instead (`scores`
/// unused; fo->count zeroed by the caller).
void snr_boxcar_3d(const float* folds, uint64_t nprof, uint32_t nbins, const WidthTable& wt,
                   float* scores, bool find_max, hipStream_t s, const FilterOut* fo = nullptr);

/// Fused inverse real DFT (matrix cores) + boxcar S/N of final Fourier fold
/// rows [ntrials, 2, nbins/2+1] float2 -> scores [ntrials, nw] or [ntrials];
/// the time-domain fold is never stored.  nbins in {16, 32, 64}.
bool irfft_score_supported(uint32_t nbins);
void irfft_score(const float2* fold, uint64_t ntrials, uint32_t nbins, const WidthTable& wt,
                 float* scores, bool find_max, bool exact, hipStream_t s,
                 const FilterOut* fo = nullptr);

/// Pruning seeds from final-level FFA folds (kernels_cuda.cu:16-277): rows of
/// [e: W][v: W], W = nbins floats (time) or nbf float2 (Fourier).  ffa_rows /
/// ffa_segments bound the indices (an index beyond them gives NaN output).
void seed_linear(const float* tree, const uint32_t* itree, const uint8_t* mask, const float* ffa,
                 const uint32_t* iffa, const float* shift, float* out, uint32_t nbins,
                 uint64_t nleaves, uint32_t start, uint32_t capacity, uint64_t ffa_rows,
                 hipStream_t s);
void seed_ascend(const float* ffa, const uint32_t* iseg, const uint32_t* iffa,
                 const float* shift, float* out, uint32_t nbins, uint64_t ncoords_init,
                 uint64_t nleaves, uint32_t nsegments, uint64_t ffa_segments, hipStream_t s);
void seed_linear_complex(const float2* tree, const uint32_t* itree, const uint8_t* mask,
                         const float2* ffa, const uint32_t* iffa, const float* shift,
                         float2* out, uint32_t nbf, uint32_t nbins, uint64_t nleaves,
                         uint32_t start, uint32_t capacity, uint64_t ffa_rows, hipStream_t s);
void seed_ascend_complex(const float2* ffa, const uint32_t* iseg, const uint32_t* iffa,
                         const float* shift, float2* out, uint32_t nbf, uint32_t nbins,
                         uint64_t ncoords_init, uint64_t nleaves, uint32_t nsegments,
                         uint64_t ffa_segments, hipStream_t s);

/// snr_boxcar_2d_max: folds [nprof, nbins] (already normalised) / stdnoise.
void snr_boxcar_2d_max(const float* folds, uint64_t nprof, uint32_t nbins, const WidthTable& wt,
                       float stdnoise, float* scores, hipStream_t s);

}  // namespace loki_amd::kern
