// FFA plan: the drop-in counterpart of loki's FFAPlanBase / FFAPlan<T>
// (include/loki/common/plans.hpp:21-201, lib/plans.cpp:264-565).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "config.hpp"

namespace loki_amd {

/// Per-coordinate merge record, same 16-byte layout as coord.hpp:15-20.
struct FFACoord {
    uint32_t i_tail;
    float shift_tail;
    uint32_t i_head;
    float shift_head;
};

/// Frequency-only merge record, same 8-byte layout as coord.hpp:22-25.
struct FFACoordFreq {
    uint32_t idx;
    float shift;
};

enum class FoldKind : int { kTime = 0, kFourier = 1 };

class FFAPlan {
public:
    FFAPlan(const SearchConfig& cfg, FoldKind kind);

    const SearchConfig& config() const noexcept { return m_cfg; }
    FoldKind kind() const noexcept { return m_kind; }
    SizeType n_params() const noexcept { return m_n_params; }
    SizeType n_levels() const noexcept { return m_n_levels; }
    const std::vector<SizeType>& segment_lens() const noexcept { return m_segment_lens; }
    const std::vector<SizeType>& nsegments() const noexcept { return m_nsegments; }
    const std::vector<double>& tsegments() const noexcept { return m_tsegments; }
    const std::vector<SizeType>& ncoords() const noexcept { return m_ncoords; }
    const std::vector<float>& ncoords_lb() const noexcept { return m_ncoords_lb; }
    const std::vector<uint32_t>& ncoords_offsets() const noexcept { return m_ncoords_offsets; }
    /// 64-bit twin of ncoords_offsets (the reference's u32 offsets wrap for
    /// coordinate tables above 2^32 entries).
    const std::vector<SizeType>& ncoords_offsets64() const noexcept { return m_offsets64; }
    const std::vector<std::vector<SizeType>>& param_counts() const noexcept { return m_counts; }
    const std::vector<uint32_t>& param_counts_flat() const noexcept { return m_counts_flat; }
    const std::vector<std::vector<SizeType>>& param_cart_strides() const noexcept {
        return m_strides;
    }
    const std::vector<std::vector<double>>& dparams() const noexcept { return m_dparams; }
    const std::vector<std::vector<double>>& dparams_actual() const noexcept {
        return m_dparams_act;
    }
    const std::vector<std::vector<SizeType>>& fold_shapes() const noexcept { return m_shapes; }
    const std::vector<std::vector<SizeType>>& fold_shapes_time() const noexcept {
        return m_shapes_time;
    }

    SizeType fold_width() const noexcept;  ///< nbins (time) or nbins_f (Fourier)
    SizeType level_size(SizeType level) const;  ///< elements of fold level `level`
    SizeType coord_size() const noexcept;
    SizeType brute_fold_size() const;
    SizeType fold_size() const;
    SizeType fold_size_time() const;
    SizeType buffer_size() const;
    SizeType buffer_size_time() const;
    float buffer_memory_usage() const;
    float coord_memory_usage() const;
    float gflops(bool return_in_time) const;

    std::vector<std::vector<double>> compute_param_grid(SizeType level) const;
    std::map<std::string, std::vector<double>> params_dict() const;

private:
    SearchConfig m_cfg;
    FoldKind m_kind;
    SizeType m_n_params{};
    SizeType m_n_levels{};
    std::vector<SizeType> m_segment_lens, m_nsegments, m_ncoords, m_offsets64;
    std::vector<double> m_tsegments;
    std::vector<float> m_ncoords_lb;
    std::vector<uint32_t> m_ncoords_offsets, m_counts_flat;
    std::vector<std::vector<SizeType>> m_counts, m_strides, m_shapes, m_shapes_time;
    std::vector<std::vector<double>> m_dparams, m_dparams_act;
    SizeType m_flops{}, m_flops_time{};
};

}  // namespace loki_amd
