// Frequency sweep (ffa_freq_sweep_cuda.cu:84-342): the region planner's chunks
// run one after another on one stream, each as FFA + score-and-filter at
// cfg.snr_min into one growing device candidate list; the candidates become
// (parameters..., width) rows at the end.  Every chunk is scored with the base
// config's boxcar widths (ffa_freq_sweep_cuda.cu:64, 261), although a chunk's
// own config (more bins in a lower band) would list more.  The reference reads the passing
// count back after every chunk; so does this (it bounds the next chunk's
// output), but nothing else leaves the device until the end.
#include "sweep.hpp"

#include <algorithm>

#include "grid.hpp"

namespace loki_amd {

FFAFreqSweep::FFAFreqSweep(const SearchConfig& cfg, int device_id)
    : m_base(cfg),
      m_kind(cfg.get_use_fourier() ? FoldKind::kFourier : FoldKind::kTime),
      m_device(device_id),
      m_planner(cfg, m_kind, /*use_gpu=*/true) {
    LOKI_REQUIRE(m_planner.nregions() > 0, "FFAFreqSweep: the region planner found no chunks");
}

SizeType FFAFreqSweep::execute(const float* ts_e, const float* ts_v, SizeType nsamps) {
    DeviceGuard g(m_device);
    LOKI_REQUIRE(ts_e != nullptr && ts_v != nullptr, "FFAFreqSweep: null input");
    LOKI_REQUIRE(nsamps == m_base.get_nsamps(), "FFAFreqSweep: ts_e must have size nsamps");
    const SizeType cap = m_planner.stats().max_scores_size();
    LOKI_REQUIRE(cap < (1ULL << 32), "FFAFreqSweep: candidate buffer exceeds 32-bit indices");
    if (m_de.get() == nullptr) {
        m_de   = DeviceBuffer(nsamps * 4);
        m_dv   = DeviceBuffer(nsamps * 4);
        m_didx = DeviceBuffer(cap * 4);
        m_dsnr = DeviceBuffer(cap * 4);
        m_dcnt = DeviceBuffer(4);
    }
    DeviceBuffer &de = m_de, &dv = m_dv, &didx = m_didx, &dsnr = m_dsnr, &dcnt = m_dcnt;
    hip_check(hipMemcpy(de.get(), ts_e, nsamps * 4, hipMemcpyHostToDevice), "copy ts_e");
    hip_check(hipMemcpy(dv.get(), ts_v, nsamps * 4, hipMemcpyHostToDevice), "copy ts_v");
    const auto& widths = m_base.get_scoring_widths();
    m_per_region.assign(m_planner.nregions(), 0);
    if (m_exec.size() != m_planner.nregions()) {
        m_ws = std::make_shared<FFAWorkspace>();
        m_exec.clear();
        for (const auto& c : m_planner.cfgs()) {
            m_exec.push_back(std::make_unique<FFAExecutor>(c, m_kind, m_device, m_ws));
        }
    }
    SizeType total = 0;
    for (SizeType r = 0; r < m_planner.nregions(); ++r) {
        FFAExecutor& ffa = *m_exec[r];
        kern::FilterOut fo{};
        fo.threshold = static_cast<float>(m_base.get_snr_min());
        fo.capacity  = static_cast<uint32_t>(cap - total);
        fo.count     = dcnt.as<uint32_t>();
        fo.indices   = didx.as<uint32_t>() + total;
        fo.snr       = dsnr.as<float>() + total;
        ffa.execute_scores(de.as<float>(), dv.as<float>(), widths, nullptr, 0, false, nullptr,
                           &fo);
        uint32_t n = 0;
        hip_check(hipMemcpy(&n, dcnt.get(), 4, hipMemcpyDeviceToHost), "read candidate count");
        LOKI_REQUIRE(n <= cap - total,
                     "FFAFreqSweep: candidate buffer full; raise snr_min or "
                     "max_passing_candidates");
        m_per_region[r] = n;
        total += n;
    }
    m_idx.resize(total);
    m_snr.resize(total);
    if (total > 0) {
        hip_check(hipMemcpy(m_idx.data(), didx.get(), total * 4, hipMemcpyDeviceToHost),
                  "copy candidate indices");
        hip_check(hipMemcpy(m_snr.data(), dsnr.get(), total * 4, hipMemcpyDeviceToHost),
                  "copy candidate scores");
    }
    build_rows();
    return total;
}

// ffa_freq_sweep_cuda.cu:263-340: index -> (trial, width) -> grid values
// (cell centres of the chunk's final grid, psr_utils.hpp:283-293) + width.
// Within a chunk the rows are put in index order (the device order depends on
// scheduling).
void FFAFreqSweep::build_rows() {
    const SizeType np   = m_base.get_nparams();
    const auto& widths  = m_base.get_scoring_widths();
    const SizeType nw   = widths.size();
    const SizeType total = m_snr.size();
    m_rows.assign(total * (np + 1), 0.0);
    std::vector<float> snr_sorted(total);
    SizeType off = 0;
    for (SizeType r = 0; r < m_planner.nregions(); ++r) {
        const SearchConfig& cfg = m_planner.cfgs()[r];
        const FFAPlan plan(cfg, m_kind);
        const auto& counts = plan.param_counts().back();
        const auto lim     = cfg.get_param_limits();
        const SizeType n   = m_per_region[r];
        std::vector<SizeType> order(n);
        for (SizeType i = 0; i < n; ++i) {
            order[i] = i;
        }
        std::sort(order.begin(), order.end(),
                  [&](SizeType a, SizeType b) { return m_idx[off + a] < m_idx[off + b]; });
        for (SizeType k = 0; k < n; ++k) {
            const SizeType src  = off + order[k];
            const SizeType dst  = off + k;
            SizeType coord      = m_idx[src] / nw;
            const SizeType w    = m_idx[src] % nw;
            double* row         = m_rows.data() + dst * (np + 1);
            for (SizeType j = np; j-- > 0;) {  // frequency fastest
                const SizeType c = counts[j];
                const SizeType i = coord % c;
                coord /= c;
                row[j] = grid::param_val_at_idx(lim[j].min, lim[j].max, static_cast<uint32_t>(c),
                                                static_cast<uint32_t>(i));
            }
            row[np]          = static_cast<double>(widths[w]);
            snr_sorted[dst]  = m_snr[src];
        }
        off += n;
    }
    m_snr = std::move(snr_sorted);
}

}  // namespace loki_amd
