/*
 * loki_amd.h -- C-ABI of the MI355X (gfx950) fast-folding (FFA) hot path.
 *
 * Drop-in boundary for pravirkr/loki's CUDA backend (libculoki).  Every entry
 * point below names the reference interface it replaces (file:line in
 * pravirkr/loki v0.0.3).  Plain pointers and sizes only; no C++ or torch
 * types.  All functions return LOKI_OK (0) or an error code; the message of
 * the last failure on the calling thread is returned by loki_last_error().
 *
 * Device entry points (*_device) take device pointers and a hipStream_t
 * passed as void*; they are asynchronous on that stream (the reference
 * synchronises per level: kernels_cuda.cu:1081-1082 -- this library does not).
 * Host entry points (*_host) take host pointers and are synchronous, like the
 * reference's host-span overloads.
 */
#ifndef LOKI_AMD_H
#define LOKI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LOKI_OK 0
#define LOKI_ERR_INVALID 1  /* bad argument / config (reference: DetailedException) */
#define LOKI_ERR_HIP 2      /* HIP runtime failure (reference: CudaException)     */
#define LOKI_ERR_INTERNAL 3

#define LOKI_FOLD_TIME 0
#define LOKI_FOLD_FOURIER 1

typedef struct loki_config_s* loki_config_t;
typedef struct loki_plan_s* loki_plan_t;
typedef struct loki_ffa_s* loki_ffa_t;
typedef struct loki_batch_s* loki_batch_t;

/* PulsarSearchConfig constructor arguments (configs.hpp:25-47).
 * param_limits: nparams (min, max) pairs, order (..., jerk, accel, freq).
 * bseg_brute / bseg_ffa: 0 selects the reference default (configs.cpp:295-302). */
typedef struct {
    size_t nsamps;
    double tsamp;
    size_t nbins;
    double eta;
    const double* param_limits;
    size_t nparams;
    double ducy_max;
    double wtsp;
    int use_fourier;
    int nthreads;
    double max_process_memory_gb;
    double octave_scale;
    size_t nbins_max;
    size_t nbins_min_lossy_bf;
    size_t bseg_brute;
    size_t bseg_ffa;
    double snr_min;
    size_t max_passing_candidates;
    size_t prune_poly_order;
    double p_orb_min;
    double m_c_max;
    double m_p_min;
    double minimum_snap_cells;
    int use_conservative_tile;
} loki_config_params;

/* Derived config scalars (configs.cpp:77-83, getters configs.hpp:57-128). */
typedef struct {
    size_t nbins_f;
    size_t nparams;
    size_t bseg_brute;
    size_t bseg_ffa;
    size_t niters_ffa;
    size_t n_scoring_widths;
    double tobs;
    double tseg_brute;
    double tseg_ffa;
    double f_min;
    double f_max;
} loki_config_info;

/* FFAPlan sizes (plans.hpp:36-190). */
typedef struct {
    size_t n_params;
    size_t n_levels;
    size_t brute_fold_size;
    size_t fold_size;
    size_t fold_size_time;
    size_t buffer_size;
    size_t buffer_size_time;
    size_t coord_size;
    float buffer_memory_usage;
    float coord_memory_usage;
    float gflops;
    float gflops_return_in_time;
} loki_plan_info;

/* One FFAPlan level: segment_lens / nsegments / tsegments / ncoords /
 * ncoords_offsets / param_counts / dparams / dparams_actual / fold_shapes
 * entries of that level (plans.cpp:280-320, 491-518). */
typedef struct {
    size_t segment_len;
    size_t nsegments;
    double tsegment;
    size_t ncoords;
    size_t ncoords_offset;
    float ncoords_lb;
    size_t param_counts[5];
    size_t param_cart_strides[5];
    double dparams[5];
    double dparams_actual[5];
    size_t fold_shape[8];
    size_t fold_shape_time[8];
    size_t fold_ndim;
} loki_plan_level;

/* Message of the last failure on this thread ("" if none). */
const char* loki_last_error(void);
/* Number of visible HIP devices. */
int loki_device_count(int* count);

/* --- PulsarSearchConfig (configs.hpp:22-164) --------------------------- */
int loki_config_create(const loki_config_params* params, loki_config_t* out);
void loki_config_destroy(loki_config_t cfg);
int loki_config_info_get(loki_config_t cfg, loki_config_info* out);
/* Copies the constructor arguments back (param_limits points into cfg). */
int loki_config_params_get(loki_config_t cfg, loki_config_params* out);
/* score_widths (configs.cpp:82-83); *n receives the count. */
int loki_config_scoring_widths(loki_config_t cfg, size_t* out, size_t cap, size_t* n);
/* kind 0: get_dparams, 1: get_dparams_f, 2: get_dparams_actual
 * (configs.cpp:152-174); out has nparams entries. */
int loki_config_dparams(loki_config_t cfg, double tseg_cur, int kind, double* out);
/* get_param_grid_count (configs.cpp:175-184). */
int loki_config_param_grid_count(loki_config_t cfg, double tseg_cur, size_t* out);
/* generate_box_width_trials (score.cpp:390-405). */
int loki_generate_box_width_trials(size_t nbins, double ducy_max, double wtsp, size_t* out,
                                   size_t cap, size_t* n);

/* --- FFAPlan (plans.hpp:21-201, FFAPlan<float> / FFAPlan<ComplexType>) -- */
int loki_plan_create(loki_config_t cfg, int fold_type, loki_plan_t* out);
void loki_plan_destroy(loki_plan_t plan);
int loki_plan_info_get(loki_plan_t plan, loki_plan_info* out);
int loki_plan_level_get(loki_plan_t plan, size_t level, loki_plan_level* out);
/* compute_param_grid(level)[param] (plans.cpp:152-162); *n receives count. */
int loki_plan_param_grid(loki_plan_t plan, size_t level, size_t param, double* out, size_t cap,
                         size_t* n);
/* resolve_coordinates[_freq] (plans.cpp:64-150), evaluated by the FP64 HIP
 * resolver on device_id: coord_size records of 16 B {u32 i_tail, f32
 * shift_tail, u32 i_head, f32 shift_head} (nparams >= 2) or 8 B {u32 idx,
 * f32 shift} (nparams == 1), levels concatenated at ncoords_offsets, level 0
 * zero-filled.  Host output. */
int loki_plan_resolve_coordinates_host(loki_plan_t plan, void* coords, size_t coords_bytes,
                                       int device_id);

/* --- FFA executor (ffa.hpp:100-188 FFACUDA<T>) ------------------------- */
int loki_ffa_create(loki_config_t cfg, int fold_type, int device_id, loki_ffa_t* out);
/* workspace.hpp:169-206 FFAWorkspaceCUDA and the FFACUDA(workspace, cfg, device_id) constructor
 * (ffa.hpp:100-153): FFAs made with one workspace share its device buffers (the ping-pong
 * partner fold, the second scores buffer and the shift tables), each growing them to its plan,
 * so many plans (the chunks of a frequency sweep) hold one set of buffers sized by the largest.
 * They must not run concurrently.  buffer_bytes / coord_bytes (0: none) preallocate.  The
 * workspace lives until it and every FFA made with it are destroyed. */
typedef struct loki_ffa_workspace_s* loki_ffa_workspace_t;
int loki_ffa_workspace_create(int device_id, size_t buffer_bytes, size_t coord_bytes,
                              loki_ffa_workspace_t* out);
void loki_ffa_workspace_destroy(loki_ffa_workspace_t ws);
int loki_ffa_create_with_workspace(loki_config_t cfg, int fold_type, loki_ffa_workspace_t ws,
                                   int device_id, loki_ffa_t* out);
void loki_ffa_destroy(loki_ffa_t ffa);
int loki_ffa_plan(loki_ffa_t ffa, loki_plan_t* out_plan_view); /* borrowed; do not destroy */
/* FFACUDA::execute(device spans, stream) (ffa.hpp:130-136): fold holds
 * buffer_size elements; the final level is its first fold_size elements. */
int loki_ffa_execute_device(loki_ffa_t ffa, const float* ts_e, const float* ts_v, void* fold,
                            size_t fold_len, void* stream);
/* FFACUDA::execute(host spans) (ffa.hpp:125-129): synchronous. */
int loki_ffa_execute_host(loki_ffa_t ffa, const float* ts_e, const float* ts_v, size_t nsamps,
                          void* fold, size_t fold_len);
/* FFACUDA<complex>::execute(..., span<float> fold, stream) (ffa.hpp:138-148,
 * ffa_cuda.cu:248-290): Fourier FFA returned to the time domain.  fold holds
 * 2*buffer_size floats; the result is its first fold_size_time floats
 * [nseg, ncoords, 2, nbins].  Needs a Fourier (fold_type 1) object. */
int loki_ffa_execute_time_device(loki_ffa_t ffa, const float* ts_e, const float* ts_v,
                                 float* fold, size_t fold_len, void* stream);
/* The same from host spans (ffa_cuda.cu:198-246): synchronous. */
int loki_ffa_execute_time_host(loki_ffa_t ffa, const float* ts_e, const float* ts_v,
                               size_t nsamps, float* fold, size_t fold_len);
/* compute_ffa_scores_cuda (ffa_cuda.cu:670-692) on resident buffers, with
 * explicit widths (widths[i] in [1, nbins)).  scores: [ncoords_final, nw]
 * or [ncoords_final] if find_max. */
int loki_ffa_scores_device(loki_ffa_t ffa, const float* ts_e, const float* ts_v,
                           const size_t* widths, size_t nwidths, float* scores,
                           size_t scores_len, int find_max, void* stream);
int loki_ffa_scores_host(loki_ffa_t ffa, const float* ts_e, const float* ts_v, size_t nsamps,
                         const size_t* widths, size_t nwidths, float* scores, size_t scores_len,
                         int find_max);

/* Per-kernel timing of the last execute on an FFA object (hipEvent pairs recorded on the
 * launch stream around each kernel).  No reference counterpart: it replaces the reference's
 * host-side ScopeTimer / get_brute_fold_timing (ffa.hpp:52, ffa.cpp:74,204-239).
 * kind: 0 shift-table resolve, 1 brute fold, 2 merge, 3 score, 4 fused last level
 * (shift tables + merge + score), 5 real FFT between time and Fourier folds, 6 fused inverse
 * real DFT + score of the last Fourier level; level: FFA level. */
typedef struct {
    int kind;
    int level;
    float ms;
} loki_kernel_timing;
int loki_ffa_set_timing(loki_ffa_t ffa, int enable);
/* Scores path: 1 (default) fuses the last level's shift tables, merge and S/N into one kernel
 * (the final fold cube is never stored) for nbins 8, 16, 32, 64 and 128 (other nbins run
 * unfused); 0 runs the unfused merge + snr_boxcar_3d.  Both give identical results.  For Fourier plans with nbins 16, 32 or 64, 1 fuses the inverse real DFT
 * of the last level with the S/N (the time-domain final fold is never stored); both orders
 * stay within the Fourier S/N bound, not bitwise equal.  No reference counterpart (the
 * reference always stores the final fold, ffa_cuda.cu:670-692). */
int loki_ffa_set_fused(loki_ffa_t ffa, int enable);
/* Levels fused into the scores kernel of the time-domain path (nbins 8, 16 or 32): 1 the last
 * level only (its parent level is stored in HBM), 2 (default) the parent level too, merged on
 * chip from the level below it and never stored.  Same bits either way; other nbins keep the
 * one-level fused kernel.  No reference counterpart (ffa_cuda.cu:650-692 stores every level). */
int loki_ffa_set_fuse_depth(loki_ffa_t ffa, int depth);
/* Time-domain tree levels merged two at a time where the storage allows (default 1): one pass
 * reads four rows of level l - 1 per output row of level l + 1 and never stores level l.  Same
 * bits as one pass per level (0).  No reference counterpart (kernels_cuda.cu:280-383 merges one
 * level per launch). */
int loki_ffa_set_two_level_merges(loki_ffa_t ffa, int enable);
/* Profile normalisation in the fused scores path: 1 (default) the CPU's correctly rounded
 * e * (1/sqrt(v)), bit-identical to the reference CPU; 0 the hardware reciprocal square root
 * (|dS/N| ~ 1e-6, the analogue of the reference CUDA path's fast-math e/sqrtf(v),
 * score_cuda.cu:82).  No reference counterpart as a switch. */
int loki_ffa_set_exact_norm(loki_ffa_t ffa, int exact);
/* Fourier brute fold precision: 0 (default) FP32 matrix cores; 1 bf16 matrix cores with each
 * operand split into high and low bf16 parts (three products per term, FP32 accumulation,
 * ~2^-17 relative per product) for nbins <= 129.  No reference counterpart (the reference's
 * Fourier brute fold is FP32, kernels_cuda.cu:780-965). */
int loki_ffa_set_dft_precision(loki_ffa_t ffa, int mode);
/* Storage precision of the intermediate fold levels 0..L-2 of a time-domain FFA object:
 * LOKI_STORAGE_F32 (default), LOKI_STORAGE_F16, LOKI_STORAGE_BF16 or LOKI_STORAGE_BFP16.  Each
 * add is done in FP32 and each stored value is rounded once (round to nearest even).  F16/BF16:
 * level l is stored at the exact scale 2^-(k0 + l), k0 = floor(log2(segment_len / nbins)); FP16
 * assumes inputs of order unity per sample (bucket sums beyond 65504 * 2^k0 overflow).  BFP16
 * (block floating point; power-of-two nbins in [4, 256], finite inputs): levels 1..L-2 store
 * each component row as nbins int16 mantissas with one shared exponent chosen from the row's
 * largest magnitude (15 significant bits relative to it), level 0 stays FP32.  The final level
 * is returned and scored in FP32.  No reference counterpart (the reference stores FP32 or
 * complex64 folds, plans.hpp:151-201). */
#define LOKI_STORAGE_F32 0
#define LOKI_STORAGE_F16 1
#define LOKI_STORAGE_BF16 2
#define LOKI_STORAGE_BFP16 3
int loki_ffa_set_fold_storage(loki_ffa_t ffa, int storage);
/* Waits for the recorded events; *n receives the number of timed launches. */
int loki_ffa_timings(loki_ffa_t ffa, loki_kernel_timing* out, size_t cap, size_t* n);

/* Score and filter (score_cuda.cu:382-443 score_and_filter_cuda_d / _max_cuda_d, used by the
 * FFAFreqSweepCUDA pipeline, ffa_freq_sweep_cuda.cu:240-257): the FFA's final-level S/N, kept
 * only where it exceeds `threshold`.  Per width (find_max 0): every (trial, width) with
 * snr > threshold, index = trial * nwidths + width; find_max 1: every trial whose best S/N
 * exceeds it, index = trial.  The final fold and the full score array are never stored (the
 * last level is merged, scored and compacted in one kernel where nbins allows).  Device form:
 * indices/scores/count are device buffers, `count` (uint32) is zeroed on `stream` and receives
 * the number that passed, which may exceed `capacity` (only `capacity` entries are written,
 * in no fixed order).  Host form: synchronous, results sorted by index, *count as above.
 * trial * nwidths must fit in 32 bits (the reference's uint32 indices). */
int loki_ffa_scores_filter_device(loki_ffa_t ffa, const float* ts_e, const float* ts_v,
                                  const size_t* widths, size_t nwidths, float threshold,
                                  int find_max, uint32_t* indices, float* scores, size_t capacity,
                                  uint32_t* count, void* stream);
int loki_ffa_scores_filter_host(loki_ffa_t ffa, const float* ts_e, const float* ts_v,
                                size_t nsamps, const size_t* widths, size_t nwidths,
                                float threshold, int find_max, uint32_t* indices, float* scores,
                                size_t capacity, size_t* count);

/* --- Batched multi-series FFA ------------------------------------------
 * No reference counterpart: the reference searches series one FFACUDA object at a time, with
 * host synchronisations per level (ffa_freq_sweep_cuda.cu:122-136, kernels_cuda.cu:1081-1082).
 * A batch holds many independent frequency-only time-domain searches (uneven lengths, grids
 * and nbins in {8, 16, 32, 64, 128}) and advances all of them with one grouped launch per
 * stage and nbins class, captured once in a HIP graph.  Each series' scores are bit-identical
 * to loki_ffa_scores_* on its own config.  Inputs are copied into the batch's own device slots
 * (loki_batch_series) before loki_batch_execute. */
typedef struct {
    uint32_t merge_group;    /* child rows of a merge scheduled back to back */
    uint32_t merge_run;      /* frequency tiles per sequential run */
    uint32_t final_group;    /* the same for the fused last levels */
    uint32_t final_run;
    uint32_t persistent;     /* must be 0: the one-launch persistent mode is disabled */
    uint32_t score_lanes[5]; /* boxcar-width schedule of the last levels, nbins 8, 16, 32, 64,
                                128: 0 = one lane per trial (64-lane workgroups); 1, 2 or 4 =
                                that many lanes per trial in 256-lane workgroups, each taking
                                every score_lanes-th width (nbins >= 4 * lanes; above 32 bins
                                nbins <= 32 * lanes) */
} loki_batch_schedule;
/* widths: boxcar widths for every series (those >= nbins dropped); NULL/0: each config's own
 * scoring widths. */
int loki_batch_create(const loki_config_t* cfgs, size_t nseries, const size_t* widths,
                      size_t nwidths, int device_id, loki_batch_t* out);
void loki_batch_destroy(loki_batch_t batch);
int loki_batch_info(loki_batch_t batch, size_t* nseries, size_t* total_trials,
                    size_t* device_bytes, size_t* launches);
int loki_batch_plan(loki_batch_t batch, size_t i, loki_plan_t* out_plan_view); /* borrowed */
/* Device slots of series i: ts_e / ts_v (nsamps floats each, to be filled by the caller) and
 * scores [ncoords_final, nwidths]. */
int loki_batch_series(loki_batch_t batch, size_t i, float** ts_e, float** ts_v, float** scores,
                      size_t* scores_len, size_t* nwidths);
int loki_batch_widths(loki_batch_t batch, size_t i, size_t* out, size_t cap, size_t* n);
/* Host copies in and out of series i's slots (synchronous; scores wait for the device). */
int loki_batch_set_input_host(loki_batch_t batch, size_t i, const float* ts_e,
                              const float* ts_v, size_t nsamps);
int loki_batch_scores_host(loki_batch_t batch, size_t i, float* out, size_t len);
/* Device-to-device copy of resident series into slot i, asynchronous on `stream`. */
int loki_batch_set_input_device(loki_batch_t batch, size_t i, const float* ts_e,
                                const float* ts_v, size_t nsamps, void* stream);
/* Runs every series on `stream` (asynchronous): replays the captured graph (default) or
 * enqueues the grouped launches directly. */
int loki_batch_execute(loki_batch_t batch, void* stream);
int loki_batch_set_graph(loki_batch_t batch, int enable);
int loki_batch_set_schedule(loki_batch_t batch, const loki_batch_schedule* schedule);
/* Times candidate schedules (tile orders, then lanes per trial of the last levels per nbins)
 * on the batch itself (synchronous) and keeps the fastest. */
int loki_batch_autotune(loki_batch_t batch, void* stream, int repeats,
                        loki_batch_schedule* chosen);
/* Waits for the device; *flags = nonzero if a launch since the last call flagged an error
 * (its scores are then invalid).  Clears the flag. */
int loki_batch_errors(loki_batch_t batch, uint32_t* flags);

/* --- Pruning seeds from final-level FFA folds ---------------------------- */
/* Rows are [e: W][v: W]: W = nbins floats (time) or nbins_f = nbins/2+1 complex64 (Fourier,
 * passed as void*).  Asynchronous on `stream`; every pointer is device memory.  Beyond the
 * reference's arguments each entry takes the number of rows (or segments) of folds_ffa: an
 * index at or beyond it, or a tree index still >= capacity after one wrap, writes NaN to that
 * leaf's bins instead of reading out of bounds. */
/* kernels.hpp:225-236 shift_add_linear_batch_cuda (kernels_cuda.cu:16-65, 1236-1261):
 * for each leaf l with validation_mask[l] != 0 (NULL: every leaf),
 *   folds_out[l] = folds_tree[(indices_tree[l] + physical_start_idx) wrapped at capacity]
 *                + rotate_right(folds_ffa[indices_ffa[l]], u32(phase_shift[l] + 0.5)). */
int loki_shift_add_linear_batch_device(const float* folds_tree, const uint32_t* indices_tree,
                                       const uint8_t* validation_mask, const float* folds_ffa,
                                       const uint32_t* indices_ffa, const float* phase_shift,
                                       float* folds_out, size_t nbins, size_t n_leaves,
                                       uint32_t physical_start_idx, uint32_t capacity,
                                       size_t ffa_rows, void* stream);
/* kernels.hpp:238-251 shift_add_linear_complex_batch_cuda (kernels_cuda.cu:141-203): as above
 * with the head multiplied by exp(-2 pi i k phase_shift[l] / nbins) per harmonic k. */
int loki_shift_add_linear_complex_batch_device(
    const void* folds_tree, const uint32_t* indices_tree, const uint8_t* validation_mask,
    const void* folds_ffa, const uint32_t* indices_ffa, const float* phase_shift,
    void* folds_out, size_t nbins_f, size_t nbins, size_t n_leaves, uint32_t physical_start_idx,
    uint32_t capacity, size_t ffa_rows, void* stream);
/* kernels.hpp:253-263 shift_add_ascend_linear_batch_cuda (kernels_cuda.cu:67-139): for every
 * leaf l, folds_tree[l] = sum over g < n_segments, in order, of
 *   rotate_right(folds_ffa[indices_segment[g]][indices_ffa[g * n_leaves + l]],
 *                u32(phase_shift[g * n_leaves + l] + 0.5));
 * folds_ffa is [ffa_segments][n_coords_init][2][nbins]. */
int loki_shift_add_ascend_linear_batch_device(const float* folds_ffa,
                                              const uint32_t* indices_segment,
                                              const uint32_t* indices_ffa,
                                              const float* phase_shift, float* folds_tree,
                                              size_t nbins, size_t n_coords_init,
                                              size_t n_leaves, size_t n_segments,
                                              size_t ffa_segments, void* stream);
/* kernels.hpp:265-276 shift_add_ascend_linear_complex_batch_cuda (kernels_cuda.cu:205-277). */
int loki_shift_add_ascend_linear_complex_batch_device(
    const void* folds_ffa, const uint32_t* indices_segment, const uint32_t* indices_ffa,
    const float* phase_shift, void* folds_tree, size_t nbins_f, size_t nbins,
    size_t n_coords_init, size_t n_leaves, size_t n_segments, size_t ffa_segments,
    void* stream);

/* --- Frequency sweep: region planner and driver -------------------------- */
/* regions.hpp:40-46 generate_ffa_regions (coord.hpp FFARegion): octave bands of the period
 * range [p_min, p_max] with nbins growing so that one bin keeps its width (p_min / nbins_min)
 * and eta growing with nbins; nbins capped at nbins_max (0: no cap).  out may be NULL to query
 * *n. */
typedef struct {
    double f_start;
    double f_end;
    size_t nbins;
    double eta;
} loki_ffa_region;
int loki_generate_ffa_regions(double p_min, double p_max, double tsamp, size_t nbins_min,
                              double eta_min, double octave_scale, size_t nbins_max,
                              loki_ffa_region* out, size_t cap, size_t* n);
/* regions.hpp FFARegionStats: maxima over the planned chunks and the memory they imply (GiB). */
typedef struct {
    size_t max_buffer_size;
    size_t max_coord_size;
    size_t max_ncoords;
    size_t max_ffa_levels;
    size_t max_buffer_size_time;
    size_t max_scores_size;
    size_t write_param_sets_size;
    float buffer_memory_usage;
    float coord_memory_usage;
    float extra_memory_usage;
    float freq_sweep_memory_usage;
} loki_region_stats;
int loki_region_stats_compute(size_t max_buffer_size, size_t max_coord_size, size_t max_ncoords,
                              size_t max_ffa_levels, size_t n_widths, size_t n_params,
                              size_t n_samps, size_t max_passing_candidates, int use_fourier,
                              int use_gpu, loki_region_stats* out);
/* coord.hpp FFAChunkStats. */
typedef struct {
    double nominal_f_start;
    double nominal_f_end;
    double actual_f_start;
    double actual_f_end;
    double nominal_width;
    double actual_width;
    double total_memory_gb;
    double overlap_fraction;
} loki_ffa_chunk_stats;
/* regions.cpp:163-420 FFARegionPlanner<T>(cfg): each band split into frequency chunks whose FFA
 * (with the drift of accel/jerk widening every chunk) fits cfg.max_process_memory_gb. */
typedef struct loki_region_planner_s* loki_region_planner_t;
int loki_region_planner_create(loki_config_t cfg, int fold_type, int use_gpu,
                               loki_region_planner_t* out);
void loki_region_planner_destroy(loki_region_planner_t planner);
int loki_region_planner_nregions(loki_region_planner_t planner, size_t* n);
/* A new config handle of chunk i (destroy with loki_config_destroy). */
int loki_region_planner_config(loki_region_planner_t planner, size_t i, loki_config_t* out);
int loki_region_planner_stats(loki_region_planner_t planner, loki_region_stats* out);
int loki_region_planner_chunk(loki_region_planner_t planner, size_t i,
                              loki_ffa_chunk_stats* out);
/* FFAFreqSweepCUDA (ffa_freq_sweep_cuda.cu:21-342): every planned chunk as FFA + score-and-filter
 * at cfg.snr_min on device `device_id`; the candidates (S/N and rows of the final-grid parameter
 * values plus the boxcar width, nparams + 1 doubles each) are kept for loki_freq_sweep_results,
 * chunk by chunk, in index order within a chunk.  The file output of the reference
 * (FFAResultWriter, HDF5) is left to the caller. */
typedef struct loki_freq_sweep_s* loki_freq_sweep_t;
int loki_freq_sweep_create(loki_config_t cfg, int device_id, loki_freq_sweep_t* out);
void loki_freq_sweep_destroy(loki_freq_sweep_t sweep);
int loki_freq_sweep_execute(loki_freq_sweep_t sweep, const float* ts_e, const float* ts_v,
                            size_t nsamps, size_t* ncandidates);
/* Candidates of the last execute (0 before the first). */
int loki_freq_sweep_count(loki_freq_sweep_t sweep, size_t* n);
int loki_freq_sweep_results(loki_freq_sweep_t sweep, float* snr, double* param_sets, size_t cap);
/* Candidates of each chunk in the last execute (ffa_freq_sweep_cuda.cu:128
 * m_n_passing_scores_per_region); `cap` >= the planner's nregions. */
int loki_freq_sweep_per_region(loki_freq_sweep_t sweep, size_t* counts, size_t cap);
/* A new planner handle with the sweep's chunks (destroy with loki_region_planner_destroy). */
int loki_freq_sweep_planner(loki_freq_sweep_t sweep, loki_region_planner_t* out);

/* --- Brute fold (fold.hpp:78-133 BruteFoldCUDA) ------------------------ */
/* compute_brute_fold_cuda<float> (fold_cuda.cu:260-277): fold has
 * (nsamps/segment_len) * nfreqs * 2 * nbins floats. */
int loki_brute_fold_time_host(const float* ts_e, const float* ts_v, size_t nsamps,
                              const double* freqs, size_t nfreqs, size_t segment_len,
                              size_t nbins, double tsamp, double t_ref, float* fold,
                              int device_id);
int loki_brute_fold_time_device(const float* ts_e, const float* ts_v, size_t nsamps,
                                const double* freqs_host, size_t nfreqs, size_t segment_len,
                                size_t nbins, double tsamp, double t_ref, float* fold,
                                void* stream);

/* No counterpart in the reference: loki_brute_fold_time_device for callers that cannot allocate
 * device memory (code that must not call hipMalloc/hipFree, such as tight serving loops or
 * harnesses that forbid allocation).  Same arguments and results (bit-identical fold); the
 * frequencies, bucket counts, offsets and lists live in a 64 MiB file-scope device array of the
 * library instead of allocated buffers (the same (freq, bin) sample lists sized by the largest
 * bucket per frequency: about nfreqs * segment_len * 4 bytes).  Returns LOKI_ERR_INVALID when
 * they need more than 64 MiB.  Each GPU holds one copy of the array, so calls that overlap on
 * one GPU race; each call synchronises `stream` before returning. */
int loki_brute_fold_time_device_noalloc(const float* ts_e, const float* ts_v, size_t nsamps,
                                        const double* freqs_host, size_t nfreqs,
                                        size_t segment_len, size_t nbins, double tsamp,
                                        double t_ref, float* fold, void* stream);

/* compute_brute_fold_cuda<ComplexTypeCUDA> (fold_cuda.cu:260-277, kernels_cuda.cu:1012-1061):
 * fold has (nsamps/segment_len) * nfreqs * 2 * (nbins/2+1) complex64 values
 * (interleaved float pairs); nbins in [2, 256]. */
int loki_brute_fold_fourier_host(const float* ts_e, const float* ts_v, size_t nsamps,
                                 const double* freqs, size_t nfreqs, size_t segment_len,
                                 size_t nbins, double tsamp, double t_ref, float* fold,
                                 int device_id);
int loki_brute_fold_fourier_device(const float* ts_e, const float* ts_v, size_t nsamps,
                                   const double* freqs_host, size_t nfreqs, size_t segment_len,
                                   size_t nbins, double tsamp, double t_ref, float* fold,
                                   void* stream);

/* --- Real transforms (fft.hpp:180-262 irfft_batch_cuda / rfft_batch_cuda) --- */
/* fft_cuda.cu:357-448: batch rows of n_real/2+1 complex64 -> n_real floats, scaled 1/n_real. */
int loki_irfft_batch_device(const float* complex_in, float* real_out, size_t batch_size,
                            size_t n_real, void* stream);
/* fft_cuda.cu:282-355: batch rows of n_real floats -> n_real/2+1 complex64 (unnormalised). */
int loki_rfft_batch_device(const float* real_in, float* complex_out, size_t batch_size,
                           size_t n_real, void* stream);

/* --- Scores (score.hpp:123-165 snr_boxcar_*_cuda) ---------------------- */
/* folds [nprof, 2, nbins] -> [nprof, nw] (find_max 0) or [nprof] (1). */
int loki_snr_boxcar_3d_device(const float* folds, size_t nprof, size_t nbins,
                              const size_t* widths, size_t nwidths, float* scores, int find_max,
                              void* stream);
int loki_snr_boxcar_3d_host(const float* folds, size_t nprof, size_t nbins,
                            const size_t* widths, size_t nwidths, float* scores, int find_max,
                            int device_id);
/* score_and_filter_cuda_d / score_and_filter_max_cuda_d (score_cuda.cu:382-443) on folds
 * [nprof, 2, nbins]: the filtered outputs of loki_ffa_scores_filter_* above. */
int loki_snr_boxcar_3d_filter_device(const float* folds, size_t nprof, size_t nbins,
                                     const size_t* widths, size_t nwidths, float threshold,
                                     int find_max, uint32_t* indices, float* scores,
                                     size_t capacity, uint32_t* count, void* stream);
int loki_snr_boxcar_3d_filter_host(const float* folds, size_t nprof, size_t nbins,
                                   const size_t* widths, size_t nwidths, float threshold,
                                   int find_max, uint32_t* indices, float* scores,
                                   size_t capacity, size_t* count, int device_id);
/* folds [nprof, nbins] -> [nprof] max over widths of snr / stdnoise. */
int loki_snr_boxcar_2d_max_host(const float* folds, size_t nprof, size_t nbins,
                                const size_t* widths, size_t nwidths, float stdnoise,
                                float* scores, int device_id);

/* Self-tests of exact scoring's normalisation (no counterpart in the reference, whose CPU path
 * computes 1.0F / std::sqrt(v) with IEEE operations, score.cpp:139-141).  The exact-mode kernels
 * replace sqrtf and the division by short correctly rounded sequences inside their ranges
 * (v in [2^-96, FLT_MAX] for the root, s in [2^-48, 2^64) for the reciprocal).  The sweep checks
 * both against the device's IEEE sqrtf and 1.0f / s for every float bit pattern in
 * [bits_lo, bits_hi): counts[0] / counts[1] = sqrt / reciprocal mismatches, counts[2] /
 * counts[3] = patterns checked, first_bad[0] / [1] = lowest mismatching pattern (0xFFFFFFFF if
 * none).  counts has 4 entries, first_bad 2.  Synchronous. */
int loki_selftest_exact_norm_sweep(int device_id, uint32_t bits_lo, uint32_t bits_hi,
                                   uint64_t* counts, uint32_t* first_bad);
/* out[4i..4i+3] = device sqrtf(x), 1.0f / x, short root, short reciprocal of x = bits[i], so
 * the device IEEE operations can be compared with the host's.  Synchronous. */
int loki_selftest_exact_norm_values(int device_id, const uint32_t* bits, size_t n, float* out);

/* --- Differentiable candidate refinement -------------------------------- */
/* No counterpart in the reference: a continuous-parameter generalisation of the Fourier brute
 * fold (kernels_cuda.cu:780-853, one segment spanning the series) with the Taylor phase model of
 * transforms.cpp:101-120, the inverse real DFT of fft_cuda.cu:357-448 and the fixed-width boxcar
 * S/N of score.cpp:95-172, plus its adjoint.  theta: ncand x (f [Hz], fdot [Hz/s],
 * fddot [Hz/s^2]) FP64 on the device, phase phi = f tau + fdot tau^2/2 + fddot tau^3/6 with
 * tau = k tsamp - t_ref.  nbins in {16, 32, 64, 128}; width in [1, nbins); 1 <= ncand < 65536.
 * All pointers are device pointers; calls are asynchronous on `stream`. */

/* Bytes of device workspace either call needs. */
size_t loki_refine_workspace_bytes(size_t nsamps, size_t ncand, size_t nbins);
/* scores [ncand] float; adj_state (may be NULL) [ncand][2][nbins/2+1] complex64: the adjoint
 * seed dS/dE_m, dS/dV_m that loki_refine_grad_device consumes. */
int loki_refine_score_device(const float* ts_e, const float* ts_v, size_t nsamps, double tsamp,
                             double t_ref, const double* theta, size_t ncand, size_t nbins,
                             size_t width, float* scores, float* adj_state, void* workspace,
                             size_t workspace_bytes, void* stream);
/* grad [ncand][3] FP64 = dS/d(f, fdot, fddot) at theta, from the adj_state of a score call at
 * the same theta (the fold is recomputed, not stored).  The boxcar position is held at the
 * forward's maximum. */
int loki_refine_grad_device(const float* ts_e, const float* ts_v, size_t nsamps, double tsamp,
                            double t_ref, const double* theta, size_t ncand, size_t nbins,
                            const float* adj_state, double* grad, void* workspace,
                            size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LOKI_AMD_H */
