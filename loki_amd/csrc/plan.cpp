// FFAPlan restatement: lib/plans.cpp:264-343 (configure/validate/strides),
// 445-565 (sizes, fold shapes, FLOP model), 152-189 (param grid, params_dict).
#include "plan.hpp"

#include <algorithm>
#include <cmath>

namespace loki_amd {

FFAPlan::FFAPlan(const SearchConfig& cfg, FoldKind kind) : m_cfg(cfg), m_kind(kind) {
    const SizeType levels = m_cfg.get_niters_ffa() + 1;
    m_n_params            = m_cfg.get_nparams();
    m_n_levels            = levels;
    LOKI_REQUIRE(levels >= 2, "FFA needs at least two levels (bseg_ffa > bseg_brute)");
    m_segment_lens.resize(levels);
    m_nsegments.resize(levels);
    m_tsegments.resize(levels);
    m_ncoords.resize(levels);
    m_ncoords_lb.resize(levels);
    m_ncoords_offsets.resize(levels + 1);
    m_offsets64.resize(levels + 1);
    m_counts_flat.resize(levels * m_n_params);
    m_counts.resize(levels);
    m_strides.resize(levels);
    m_dparams.resize(levels);
    m_dparams_act.resize(levels);
    for (SizeType l = 0; l < levels; ++l) {
        const SizeType seglen = m_cfg.get_bseg_brute() * (SizeType{1} << l);
        const double tseg     = static_cast<double>(seglen) * m_cfg.get_tsamp();
        m_segment_lens[l]     = seglen;
        m_nsegments[l]        = m_cfg.get_nsamps() / seglen;
        m_tsegments[l]        = tseg;
        m_dparams[l]          = m_cfg.get_dparams(tseg);
        m_dparams_act[l]      = m_cfg.get_dparams_actual(tseg);
        m_counts[l]           = m_cfg.get_param_grid_count(tseg);
        SizeType n            = 1;
        for (SizeType p = 0; p < m_n_params; ++p) {
            n *= m_counts[l][p];
            m_counts_flat[l * m_n_params + p] = static_cast<uint32_t>(m_counts[l][p]);
        }
        m_strides[l].assign(m_n_params, 1);
        for (int p = static_cast<int>(m_n_params) - 2; p >= 0; --p) {
            m_strides[l][p] = m_strides[l][p + 1] * m_counts[l][p + 1];
        }
        m_ncoords[l]    = n;
        m_ncoords_lb[l] = n > 0 ? std::log2(static_cast<float>(n)) : 0.0F;
        m_offsets64[l + 1] = m_offsets64[l] + n;
    }
    for (SizeType l = 0; l <= levels; ++l) {
        m_ncoords_offsets[l] = static_cast<uint32_t>(m_offsets64[l]);
    }
    // plans.cpp:322-333: only the frequency may have > 1 grid point at level 0.
    for (SizeType p = 0; p + 1 < m_n_params; ++p) {
        LOKI_REQUIRE(m_counts[0][p] == 1,
                     "FFAPlan::validate_plan: Only one parameter for higher order derivatives "
                     "is supported for the initialization for the first level");
    }
    const SizeType nbins   = m_cfg.get_nbins();
    const SizeType nbins_f = m_cfg.get_nbins_f();
    m_shapes.resize(levels);
    m_shapes_time.resize(levels);
    for (SizeType l = 0; l < levels; ++l) {
        m_shapes[l].push_back(m_nsegments[l]);
        m_shapes_time[l].push_back(m_nsegments[l]);
        for (SizeType p = 0; p < m_n_params; ++p) {
            m_shapes[l].push_back(m_counts[l][p]);
            m_shapes_time[l].push_back(m_counts[l][p]);
        }
        m_shapes[l].push_back(2);
        m_shapes_time[l].push_back(2);
        m_shapes[l].push_back(m_kind == FoldKind::kFourier ? nbins_f : nbins);
        m_shapes_time[l].push_back(nbins);
    }
    // plans.cpp:520-565 FLOP model
    const SizeType nsamps = m_cfg.get_nsamps();
    const SizeType nfreqs = m_ncoords[0];
    SizeType fb = 0, ff = 0;
    if (m_kind == FoldKind::kFourier) {
        if (nbins > m_cfg.get_nbins_min_lossy_bf()) {
            fb = (nsamps * 2 * nfreqs) +
                 static_cast<SizeType>(static_cast<double>(nfreqs * nbins) *
                                       std::log2(static_cast<double>(nbins)));
        } else {
            fb = 8 * nsamps * 2 * nfreqs * nbins_f;
        }
        for (SizeType l = 1; l < levels; ++l) {
            ff += m_ncoords[l] * m_nsegments[l] * 2 * nbins_f * 8;
        }
    } else {
        fb = nsamps * 2 * nfreqs;
        for (SizeType l = 1; l < levels; ++l) {
            ff += m_ncoords[l] * m_nsegments[l] * 2 * nbins;
        }
    }
    m_flops = fb + ff;
    if (m_kind == FoldKind::kFourier) {
        const SizeType nfft = fold_size_time() / nbins;
        m_flops_time = m_flops + static_cast<SizeType>(static_cast<double>(nfft * nbins) *
                                                       std::log2(static_cast<double>(nbins)));
    } else {
        m_flops_time = m_flops;
    }
    LOKI_REQUIRE(buffer_size() > 0, "FFAPlan::validate_plan: buffer size is zero");
}

SizeType FFAPlan::fold_width() const noexcept {
    return m_kind == FoldKind::kFourier ? m_cfg.get_nbins_f() : m_cfg.get_nbins();
}

SizeType FFAPlan::level_size(SizeType level) const {
    LOKI_REQUIRE(level < m_n_levels, "level out of range");
    SizeType n = 1;
    for (auto s : m_shapes[level]) {
        n *= s;
    }
    return n;
}

SizeType FFAPlan::coord_size() const noexcept { return m_offsets64.back(); }
SizeType FFAPlan::brute_fold_size() const { return level_size(0); }
SizeType FFAPlan::fold_size() const { return level_size(m_n_levels - 1); }

SizeType FFAPlan::fold_size_time() const {
    SizeType n = 1;
    for (auto s : m_shapes_time.back()) {
        n *= s;
    }
    return n;
}

SizeType FFAPlan::buffer_size() const {
    SizeType m = 0;
    for (SizeType l = 0; l < m_n_levels; ++l) {
        m = std::max(m, level_size(l));
    }
    return m;
}

SizeType FFAPlan::buffer_size_time() const {
    return m_kind == FoldKind::kFourier ? 2 * buffer_size() : buffer_size();
}

float FFAPlan::buffer_memory_usage() const {
    const SizeType elem = m_kind == FoldKind::kFourier ? 8 : 4;
    return static_cast<float>(2 * buffer_size() * elem) / static_cast<float>(1ULL << 30U);
}

float FFAPlan::coord_memory_usage() const {
    const SizeType elem = m_n_params == 1 ? sizeof(FFACoordFreq) : sizeof(FFACoord);
    return static_cast<float>(coord_size() * elem) / static_cast<float>(1ULL << 30U);
}

float FFAPlan::gflops(bool return_in_time) const {
    return static_cast<float>(return_in_time ? m_flops_time : m_flops) * 1.0e-9F;
}

std::vector<std::vector<double>> FFAPlan::compute_param_grid(SizeType level) const {
    LOKI_REQUIRE(level < m_n_levels, "level out of range");
    std::vector<std::vector<double>> out(m_n_params);
    const auto lims = m_cfg.get_param_limits();
    for (SizeType p = 0; p < m_n_params; ++p) {
        out[p] = range_param(lims[p].min, lims[p].max, m_dparams[level][p]);
    }
    return out;
}

std::map<std::string, std::vector<double>> FFAPlan::params_dict() const {
    const auto names = m_cfg.get_param_names();
    const auto grid  = compute_param_grid(m_n_levels - 1);
    std::map<std::string, std::vector<double>> out;
    for (SizeType p = 0; p < m_n_params; ++p) {
        out[names[p]] = grid[p];
    }
    return out;
}

}  // namespace loki_amd
