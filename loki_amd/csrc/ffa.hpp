// FFA executor on one GPU: the counterpart of loki's FFACUDA<T>
// (include/loki/algorithms/ffa.hpp:100-153, lib/ffa_cuda.cu:28-524) and of
// FFAWorkspaceCUDA (workspace.hpp:169-206).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <span>
#include <vector>

#include "device.hpp"
#include "kernels.hpp"
#include "plan.hpp"
#include "storage.hpp"

namespace loki_amd {

/// Kernel classes reported by the timing hooks.
enum class KernelKind : int {
    kResolve    = 0,
    kBruteFold  = 1,
    kMerge      = 2,
    kScore      = 3,
    kMergeScore = 4,  ///< fused last level (shift tables + merge + S/N)
    kTransform  = 5,  ///< real FFT between time and Fourier folds
    kTransformScore = 6,  ///< fused inverse real DFT + S/N of the last Fourier level
};

struct KernelTiming {
    KernelKind kind;
    int level;
    float ms;
};

/// hipEvent pairs recorded around each launch on the launch stream.
class LaunchTimer {
public:
    LaunchTimer() = default;
    ~LaunchTimer();
    LaunchTimer(const LaunchTimer&)            = delete;
    LaunchTimer& operator=(const LaunchTimer&) = delete;
    LaunchTimer(LaunchTimer&& o) noexcept;
    LaunchTimer& operator=(LaunchTimer&& o) noexcept = delete;
    bool enabled() const noexcept { return m_enabled; }
    void set_enabled(bool on) noexcept { m_enabled = on; }
    void reset() noexcept { m_used = 0; }
    /// Records the start event; returns the slot to close with stop().
    int start(KernelKind kind, int level, hipStream_t s);
    void stop(int slot, hipStream_t s);
    /// Waits for the recorded events and returns the elapsed times.
    std::vector<KernelTiming> collect();

private:
    struct Slot {
        hipEvent_t a{nullptr};
        hipEvent_t b{nullptr};
        KernelKind kind{};
        int level{};
    };
    bool m_enabled{false};
    std::vector<Slot> m_slots;
    size_t m_used{0};
};

/// Device buffers an FFA borrows (workspace.hpp:169-206 FFAWorkspaceCUDA): the
/// ping-pong partner, the second scores buffer and the shift tables.  FFAs
/// that share one run one at a time; each grows the buffers to its plan.
/// Uses on different streams are ordered: a use on stream s first waits for
/// an event recorded at the end of the previous use on another stream.
struct FFAWorkspace {
    FFAWorkspace() = default;
    ~FFAWorkspace();
    FFAWorkspace(const FFAWorkspace&)            = delete;
    FFAWorkspace& operator=(const FFAWorkspace&) = delete;

    DeviceBuffer internal;
    DeviceBuffer result;
    DeviceBuffer coords;

    /// Before enqueueing work that touches the buffers on `s`.
    void begin_use(hipStream_t s);
    /// After enqueueing it.
    void end_use(hipStream_t s);

private:
    hipEvent_t m_done{nullptr};
    hipStream_t m_last{nullptr};
    bool m_recorded{false};
};

class FFAExecutor {
public:
    /// `ws`: shared workspace (ffa.hpp:100-153, the FFACUDA(workspace, cfg,
    /// device_id) constructor); null gives the FFA its own.
    FFAExecutor(const SearchConfig& cfg, FoldKind kind, int device_id,
                std::shared_ptr<FFAWorkspace> ws = nullptr);

    void set_timing(bool on) noexcept { m_timer.set_enabled(on); }
    /// Per-kernel durations of the last execute()/execute_scores() call.
    std::vector<KernelTiming> timings() { return m_timer.collect(); }

    const FFAPlan& plan() const noexcept { return m_plan; }
    int device_id() const noexcept { return m_device; }

    /// Device execute (ffa_cuda.cu:171-196): ts_e/ts_v have nsamps floats;
    /// fold has plan().buffer_size() elements (float, or float2 for Fourier);
    /// the result is the first fold_size() elements.  Asynchronous on `s`.
    void execute(const float* ts_e, const float* ts_v, void* fold, SizeType fold_len,
                 hipStream_t s);

    /// Fourier FFA returned to the time domain (ffa_cuda.cu:248-290): fold
    /// has 2 * buffer_size() floats; the result is its first
    /// fold_size_time() floats, [nseg, ncoords, 2, nbins].
    void execute_time(const float* ts_e, const float* ts_v, float* fold, SizeType fold_len,
                      hipStream_t s);

    /// FFA + boxcar S/N of the final level (nsegments must be 1): scores
    /// [ncoords_final, nw] (or [ncoords_final] when find_max).  A Fourier plan
    /// returns its final level to the time domain first (ffa_cuda.cu:670-692).
    /// With `fo` the scores are filtered instead (score_cuda.cu:382-443): every
    /// (trial, width) -- or trial, with find_max -- whose S/N exceeds
    /// fo->threshold goes to the compacted list; `scores` is unused and
    /// fo->count is zeroed here.
    void execute_scores(const float* ts_e, const float* ts_v, std::span<const SizeType> widths,
                        float* scores, SizeType scores_len, bool find_max, hipStream_t s,
                        const kern::FilterOut* fo = nullptr);

    /// Resolve the shift tables of levels 1..last_level into device memory
    /// (at their ncoords_offsets positions).
    void resolve_coordinates(void* coords_dev, hipStream_t s, SizeType last_level = ~SizeType{0});

    /// Use the fused last-level kernel in execute_scores() when nbins allows
    /// (default on).  Off: unfused merge + snr_boxcar_3d (same bits).
    void set_fused(bool on) noexcept { m_fused = on; }
    /// Levels fused into the scoring kernel of execute_scores(): 1 (the last
    /// level; its parent level is stored) or 2 (the parent level is merged on
    /// chip as well and never stored; nbins 8, 16 or 32 and at least three
    /// levels).  Same bits either way.  Default 2.
    void set_fuse_depth(int depth);
    int fuse_depth() const noexcept { return m_fuse_depth; }
    /// Profile normalisation of the fused scores path: exact (default, the
    /// CPU's correctly rounded e * (1/sqrt(v))) or the hardware rsqrt.
    void set_exact_norm(bool on) noexcept { m_exact_norm = on; }
    /// Fourier brute fold on bf16 matrix cores with split (high + low) bf16
    /// operands and FP32 accumulation instead of FP32 matrix cores (default
    /// off; nbins <= 129, else FP32).
    void set_dft_split_bf16(bool on) noexcept { m_dft_split = on; }
    /// Storage precision of fold levels 0..L-2 (time domain only); the final
    /// level is always returned and scored in FP32.  Default FP32.
    void set_fold_storage(FoldStorage st);
    /// Merge pairs of tree levels in one two-level pass where allowed
    /// (default on); off: one pass per level (same bits).
    void set_two_level_merges(bool on) noexcept { m_two_level = on; }
    bool two_level_merges() const noexcept { return m_two_level; }
    FoldStorage fold_storage() const noexcept { return m_storage; }
    bool fused_scores_supported() const;

    /// Device pointer to the last shift tables resolved by execute().
    const void* coords_device() const noexcept { return m_ws->coords.get(); }
    SizeType coord_record_bytes() const noexcept;

private:
    FFAPlan m_plan;
    int m_device;
    bool m_freq_only;
    bool m_fourier;            // complex folds (float2 elements)
    bool m_lossy;              // Fourier init = time brute fold + rfft (ffa_cuda.cu:342-353)
    std::shared_ptr<FFAWorkspace> m_ws;  // ping-pong partner, scores buffer, shift tables
    DeviceBuffer m_bf_freqs;   // level-0 frequencies (FP64)
    kern::BruteFoldLists m_bf; // transposed bucket lists of the brute fold
    bool m_buckets_ready{false};
    bool m_fused{true};
    bool m_exact_norm{true};
    bool m_dft_split{false};
    FoldStorage m_storage{FoldStorage::kF32};
    bool m_two_level{true};  ///< merge pairs of levels in one pass where allowed
    int m_k0{0};  // level l of a reduced-precision tree is stored at scale 2^-(k0 + l)
    int m_fuse_depth{2};
    LaunchTimer m_timer;
    // one-lane-per-trial last level (k_final2.hip): row chunks sharing parent
    // rows, built once per plan
    bool m_f2_ready{false};
    bool m_f2_ok{false};  // parent ranges fit the LDS budget
    std::vector<kern::RowChunk> m_f2_chunks_host;
    DeviceBuffer m_f2_chunks;
    std::vector<kern::RowTerms> m_f2_rows_host;  // (tail, head) per grid row, np >= 2
    DeviceBuffer m_f2_rows;
    // resolver row terms of every level (np >= 2), built once per plan
    std::vector<kern::RowTerms> m_rt_host;
    std::vector<SizeType> m_rt_off;  // records into m_rt_host per level
    DeviceBuffer m_rt;
    bool m_rt_ready{false};
    uint32_t m_f2_rows_per_block{1};
    uint32_t m_f2_cap{0};
    uint32_t m_f2_nblocks{0};

    void prepare_final2(hipStream_t s);
    bool final2_usable() const;

    void prepare_brute_fold(hipStream_t s);
    /// Brute fold + merge levels 1..last_level; returns the buffer holding
    /// last_level (buf_even for even levels, buf_odd for odd levels).
    /// last_fp32: last_level is stored in FP32 at its true scale even when
    /// the tree uses reduced-precision storage.
    /// coords_level: shift tables are resolved up to max(last_level,
    /// coords_level).
    void* run_levels(const float* ts_e, const float* ts_v, void* buf_even, void* buf_odd,
                     SizeType last_level, bool last_fp32, hipStream_t s,
                     SizeType coords_level = 0);
    void init_level(const float* ts_e, const float* ts_v, void* out, void* scratch,
                    bool fp32_out, hipStream_t s);
    SizeType max_level_size(SizeType last_level) const;
    /// execute_scores without the workspace ordering.
    void scores_levels(const float* ts_e, const float* ts_v, std::span<const SizeType> widths,
                       float* scores, SizeType scores_len, bool find_max, hipStream_t s,
                       const kern::FilterOut* fo);
    /// Storage of level l in a run whose last level is last_level (returned
    /// in FP32 when last_fp32).
    FoldStorage stored_as(SizeType l, SizeType last_level, bool last_fp32) const;
    /// Bytes of the largest stored level of such a run (ping-pong buffer size).
    SizeType stored_bytes(SizeType last_level, bool last_fp32) const;
    SizeType elem_bytes() const noexcept { return m_fourier ? 2 * sizeof(float) : sizeof(float); }
};

/// Geometry of merge level `level` (>= 1) of a plan for the shift-table
/// resolvers (the last three parameters: jerk, accel, freq).
kern::ResolveLevel resolve_level(const FFAPlan& plan, SizeType level);

/// Standalone brute fold (fold_cuda.cu:139-164, 260-277): fold has
/// nsamps/segment_len * nfreqs * 2 * nbins floats.
void brute_fold_time_device(const float* ts_e, const float* ts_v, SizeType nsamps,
                            std::span<const double> freqs, SizeType segment_len, SizeType nbins,
                            double tsamp, double t_ref, float* fold, hipStream_t s);
/// Standalone Fourier brute fold: fold has nsamps/segment_len * nfreqs * 2 *
/// (nbins/2+1) float2 elements.  Synchronises `s`.
void brute_fold_fourier_device(const float* ts_e, const float* ts_v, SizeType nsamps,
                               std::span<const double> freqs, SizeType segment_len,
                               SizeType nbins, double tsamp, double t_ref, float2* fold,
                               hipStream_t s);

}  // namespace loki_amd
