// Region planner of a frequency sweep (regions.cpp:31-420): octave bands with
// their own nbins and eta, each split into frequency chunks whose FFA fits the
// memory limit (the workspace is sized by the maxima over all chunks), with
// the drift of the other search parameters widening every chunk.
#pragma once

#include <vector>

#include "common.hpp"
#include "config.hpp"
#include "plan.hpp"

namespace loki_amd {

/// coord.hpp FFARegion: one octave band.
struct FFARegion {
    double f_start;
    double f_end;
    SizeType nbins;
    double eta;
};

/// regions.cpp:31-76.  nbins_max = 0: no cap.
std::vector<FFARegion> generate_ffa_regions(double p_min, double p_max, double tsamp,
                                            SizeType nbins_min, double eta_min,
                                            double octave_scale, SizeType nbins_max);

/// regions.cpp:80-161: maxima over the chunks and the memory they imply (GiB).
class FFARegionStats {
public:
    FFARegionStats() = default;
    FFARegionStats(SizeType max_buffer_size, SizeType max_coord_size, SizeType max_ncoords,
                   SizeType max_ffa_levels, SizeType n_widths, SizeType n_params,
                   SizeType n_samps, SizeType max_passing_candidates, bool use_fourier,
                   bool use_gpu);
    SizeType max_buffer_size() const noexcept { return m_max_buffer_size; }
    SizeType max_coord_size() const noexcept { return m_max_coord_size; }
    SizeType max_ncoords() const noexcept { return m_max_ncoords; }
    SizeType max_ffa_levels() const noexcept { return m_max_ffa_levels; }
    SizeType max_buffer_size_time() const noexcept;
    SizeType max_scores_size() const noexcept;
    SizeType write_param_sets_size() const noexcept;
    float buffer_memory_usage() const noexcept;
    float coord_memory_usage() const noexcept;
    float extra_memory_usage() const noexcept;
    float cpu_memory_usage() const noexcept;
    float device_memory_usage() const noexcept;
    float freq_sweep_memory_usage() const noexcept;

private:
    SizeType m_max_buffer_size{0}, m_max_coord_size{0}, m_max_ncoords{0}, m_max_ffa_levels{0};
    SizeType m_n_widths{0}, m_n_params{0}, m_n_samps{0}, m_max_passing_candidates{0};
    bool m_use_fourier{false}, m_use_gpu{false};
};

/// coord.hpp FFAChunkStats.
struct FFAChunkStats {
    double nominal_f_start, nominal_f_end, actual_f_start, actual_f_end;
    double nominal_width, actual_width;
    double total_memory_gb, overlap_fraction;
};

/// regions.cpp:163-420 FFARegionPlanner<FoldType>.
class FFARegionPlanner {
public:
    FFARegionPlanner(const SearchConfig& cfg, FoldKind kind, bool use_gpu);
    const std::vector<SearchConfig>& cfgs() const noexcept { return m_cfgs; }
    SizeType nregions() const noexcept { return m_cfgs.size(); }
    const FFARegionStats& stats() const noexcept { return m_stats; }
    const std::vector<FFAChunkStats>& chunk_stats() const noexcept { return m_chunks; }

private:
    SearchConfig m_base;
    FoldKind m_kind;
    bool m_use_gpu;
    std::vector<SearchConfig> m_cfgs;
    std::vector<FFAChunkStats> m_chunks;
    FFARegionStats m_stats;

    double max_drift() const;
    void subdivide(const FFARegion& r, double drift, SizeType& mbuf, SizeType& mcoord,
                   SizeType& mnc, SizeType& mlv);
};

/// regions.hpp kFFAFreqSweepWriteBatchSize: candidate rows written per batch.
constexpr SizeType kFFAFreqSweepWriteBatchSize = 1U << 16U;

}  // namespace loki_amd
