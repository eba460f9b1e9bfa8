// Frequency sweep over the region planner's chunks (ffa_freq_sweep_cuda.cu:21-342).
#pragma once

#include <memory>
#include <vector>

#include "device.hpp"
#include "ffa.hpp"
#include "regions.hpp"

namespace loki_amd {

class FFAFreqSweep {
public:
    FFAFreqSweep(const SearchConfig& cfg, int device_id);
    /// Runs every chunk on host inputs; returns the number of candidates.
    SizeType execute(const float* ts_e, const float* ts_v, SizeType nsamps);
    const FFARegionPlanner& planner() const noexcept { return m_planner; }
    /// Candidates of the last execute(): S/N and rows (params..., width),
    /// chunk by chunk, index order within a chunk.
    const std::vector<float>& snr() const noexcept { return m_snr; }
    const std::vector<double>& rows() const noexcept { return m_rows; }
    const std::vector<SizeType>& per_region() const noexcept { return m_per_region; }

private:
    SearchConfig m_base;
    FoldKind m_kind;
    int m_device;
    FFARegionPlanner m_planner;
    /// Chunk executors, kept between calls; they share one workspace (the
    /// fold and shift-table buffers, sized by the largest chunk: the
    /// planner's memory model, ffa_freq_sweep_cuda.cu:44-66); series and
    /// candidate buffers.
    std::shared_ptr<FFAWorkspace> m_ws;
    std::vector<std::unique_ptr<FFAExecutor>> m_exec;
    DeviceBuffer m_de, m_dv, m_didx, m_dsnr, m_dcnt;
    std::vector<uint32_t> m_idx;
    std::vector<float> m_snr;
    std::vector<double> m_rows;
    std::vector<SizeType> m_per_region;

    void build_rows();
};

}  // namespace loki_amd
