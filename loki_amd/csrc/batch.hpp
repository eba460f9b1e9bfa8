// Batched FFA over many independent time series of uneven length and
// frequency range: every stage of every series is advanced by grouped launches
// (one per stage and nbins class), captured once in a HIP graph.  A
// one-launch persistent mode (k_tree.hip: tickets in dependency order,
// per-item completion counters) is built but disabled: its launches stalled
// on MI355X.  The reference runs series one FFACUDA object at a time
// with host synchronisations per level (ffa_freq_sweep_cuda.cu:122-136,
// kernels_cuda.cu:1081-1082); it has no multi-series batching.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "device.hpp"
#include "ffa.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace loki_amd {

/// Schedule knobs: tile order of the merges and last levels (rows of child
/// coordinates scheduled back to back, frequency tiles per run), the launch
/// mode, and the boxcar-width schedule of the last levels per nbins (index
/// log2(nbins) - 3): 0 is one lane per trial in 64-lane workgroups, 1, 2 or 4
/// that many lanes per trial in 256-lane workgroups, each lane taking every
/// lanes-th width.  persistent = 1 (one launch for every stage) is disabled:
/// set_schedule rejects it.
struct BatchSchedule {
    uint32_t merge_group{16};
    uint32_t merge_run{16};
    uint32_t final_group{16};
    uint32_t final_run{16};
    uint32_t persistent{0};
    uint32_t score_lanes[5]{0, 0, 0, 0, 0};
};

class FFABatch {
public:
    /// cfgs: one time-domain, frequency-only search per series (nsegments of
    /// the last level must be 1; nbins a power of two in [8, 128]).  widths:
    /// boxcar widths applied to every series (those >= nbins are dropped);
    /// empty: each series' own scoring widths.
    FFABatch(const std::vector<SearchConfig>& cfgs, const std::vector<SizeType>& widths,
             int device_id);
    ~FFABatch();
    FFABatch(const FFABatch&)            = delete;
    FFABatch& operator=(const FFABatch&) = delete;

    SizeType size() const noexcept { return m_series.size(); }
    const FFAPlan& plan(SizeType i) const { return m_series.at(i).plan; }
    /// Device input slots of series i (nsamps floats each); fill before execute().
    float* ts_e(SizeType i) const;
    float* ts_v(SizeType i) const;
    /// Device scores of series i: [ncoords_final, nwidths].
    float* scores(SizeType i) const;
    SizeType scores_len(SizeType i) const;
    const std::vector<SizeType>& widths(SizeType i) const { return m_series.at(i).widths; }
    /// Final-level trials summed over the batch.
    SizeType total_trials() const noexcept;
    SizeType device_bytes() const noexcept;
    int device_id() const noexcept { return m_device; }

    /// Runs every series (brute fold, all merges, S/N) on `s`: replays the
    /// captured graph when graphs are on (default), else enqueues the launches.
    void execute(hipStream_t s);
    void set_graph(bool on);
    /// Times candidate schedules on the batch itself and keeps the fastest.
    BatchSchedule autotune(hipStream_t s, int repeats);
    const BatchSchedule& schedule() const noexcept { return m_sched; }
    void set_schedule(const BatchSchedule& sc);
    /// Kernel launches per execute().
    SizeType launches() const noexcept { return m_sched.persistent != 0 ? 1 : m_launches.size(); }
    /// Persistent launch: polls a dependency wait may take before it gives up
    /// and flags an error.
    void set_spin_limit(uint32_t polls);
    /// Waits for the device; nonzero if a launch since the last call flagged
    /// an error (its results are then invalid).  Clears the flag.
    uint32_t errors();

private:
    struct Series {
        FFAPlan plan;
        std::vector<SizeType> widths;
        kern::BruteFoldLists lists;
        SizeType in_off{0};      ///< floats into the input arena (ts_e; ts_v follows)
        SizeType buf_off[2]{};   ///< floats into the fold arena
        SizeType coord_off{0};   ///< records into the coordinate arena
        SizeType score_off{0};   ///< floats into the score arena
        SizeType nlevels{0};
    };
    enum class Stage : int { kResolve, kFold, kMerge, kFinal };
    struct Launch {
        Stage stage;
        uint32_t nbins;   ///< merge / final: nbins; fold: 0 list path, 1 run path
        uint32_t ntasks, nblocks;
        SizeType task_off;  ///< bytes into the task arena
        uint32_t stride;    ///< final: LDS row stride; fold: the largest nbins
        uint32_t lanes;     ///< final: lanes per trial (0: 64-lane workgroups)
        SizeType lds;       ///< final with lanes: dynamic LDS bytes
    };

    int m_device;
    std::vector<Series> m_series;
    DeviceBuffer m_in, m_fold, m_coords, m_scores, m_freqs, m_tasks;
    std::vector<Launch> m_launches;
    // persistent launch: item table and task tables in m_tasks, counters in m_sync
    kern::TreeArgs m_tree{};
    uint32_t m_tree_blocks{0};
    SizeType m_tree_lds{0};
    uint32_t m_spin_limit{1U << 21};
    DeviceBuffer m_sync;
    BatchSchedule m_sched;
    bool m_use_graph{true};
    hipGraph_t m_graph{nullptr};
    hipGraphExec_t m_exec{nullptr};

    void build_tasks();
    void build_tree(std::vector<uint8_t>& blob, std::vector<SizeType>& offs);
    void enqueue(hipStream_t s) const;
    void drop_graph();
};

}  // namespace loki_amd
