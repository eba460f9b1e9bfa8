// Search configuration: the drop-in counterpart of loki's PulsarSearchConfig
// (include/loki/search/configs.hpp:22-164, lib/configs.cpp:18-303).
#pragma once

#include <optional>
#include <span>
#include <string>
#include <vector>

#include "common.hpp"

namespace loki_amd {

struct ParamLimit {
    double min;
    double max;
};

/// Constructor arguments, same names and defaults as configs.hpp:25-47.
struct SearchConfigArgs {
    SizeType nsamps{};
    double tsamp{};
    SizeType nbins{};
    double eta{};
    std::vector<ParamLimit> param_limits;
    double ducy_max                    = 0.2;
    double wtsp                        = 1.5;
    bool use_fourier                   = true;
    int nthreads                       = 1;
    double max_process_memory_gb       = 8.0;
    double octave_scale                = 2.0;
    SizeType nbins_max                 = 1024;
    SizeType nbins_min_lossy_bf        = 64;
    std::optional<SizeType> bseg_brute = std::nullopt;
    std::optional<SizeType> bseg_ffa   = std::nullopt;
    double snr_min                     = 5.0;
    SizeType max_passing_candidates    = 1U << 22U;
    SizeType prune_poly_order          = 3;
    double p_orb_min                   = 1e-5;
    double m_c_max                     = 10.0;
    double m_p_min                     = 1.4;
    double minimum_snap_cells          = 5.0;
    bool use_conservative_tile         = false;
};

/// Grid helpers of psr_utils.cpp:14-45 / psr_utils.hpp:235-270 (host).
std::vector<double> poly_taylor_step_f(SizeType nparams, double tobs, SizeType fold_bins,
                                       double tol_bins, double t_ref);
std::vector<double> poly_taylor_step_d_f(SizeType nparams, double tobs, SizeType fold_bins,
                                         double tol_bins, double f_max, double t_ref);
SizeType range_param_count(double vmin, double vmax, double dv);
std::vector<double> range_param(double vmin, double vmax, double dv);
/// score.cpp:390-405.
std::vector<SizeType> generate_box_width_trials(SizeType fold_bins, double ducy_max,
                                                double wtsp);

class SearchConfig {
public:
    explicit SearchConfig(SearchConfigArgs args);

    const SearchConfigArgs& args() const noexcept { return m_args; }
    SizeType get_nsamps() const noexcept { return m_args.nsamps; }
    double get_tsamp() const noexcept { return m_args.tsamp; }
    SizeType get_nbins() const noexcept { return m_args.nbins; }
    SizeType get_nbins_f() const noexcept { return (m_args.nbins / 2) + 1; }
    double get_tobs() const noexcept {
        return static_cast<double>(m_args.nsamps) * m_args.tsamp;
    }
    double get_eta() const noexcept { return m_args.eta; }
    std::span<const ParamLimit> get_param_limits() const noexcept {
        return m_args.param_limits;
    }
    double get_ducy_max() const noexcept { return m_args.ducy_max; }
    double get_wtsp() const noexcept { return m_args.wtsp; }
    bool get_use_fourier() const noexcept { return m_args.use_fourier; }
    int get_nthreads() const noexcept { return m_args.nthreads; }
    SizeType get_nbins_min_lossy_bf() const noexcept { return m_args.nbins_min_lossy_bf; }
    SizeType get_bseg_brute() const noexcept { return m_bseg_brute; }
    SizeType get_bseg_ffa() const noexcept { return m_bseg_ffa; }
    double get_tseg_brute() const noexcept { return m_tseg_brute; }
    double get_tseg_ffa() const noexcept { return m_tseg_ffa; }
    SizeType get_niters_ffa() const noexcept { return m_niters_ffa; }
    SizeType get_nparams() const noexcept { return m_args.param_limits.size(); }
    std::vector<std::string> get_param_names() const;
    double get_f_min() const noexcept { return m_args.param_limits.back().min; }
    double get_f_max() const noexcept { return m_args.param_limits.back().max; }
    const std::vector<SizeType>& get_scoring_widths() const noexcept { return m_widths; }

    std::vector<double> get_dparams_f(double tseg_cur) const;
    std::vector<double> get_dparams(double tseg_cur) const;
    std::vector<double> get_dparams_actual(double tseg_cur) const;
    std::vector<SizeType> get_param_grid_count(double tseg_cur) const;

    double get_max_process_memory_gb() const noexcept { return m_args.max_process_memory_gb; }
    double get_octave_scale() const noexcept { return m_args.octave_scale; }
    SizeType get_nbins_max() const noexcept { return m_args.nbins_max; }
    double get_snr_min() const noexcept { return m_args.snr_min; }
    SizeType get_max_passing_candidates() const noexcept {
        return m_args.max_passing_candidates;
    }
    /// configs.cpp:187-223 get_updated_config: the same search with another
    /// nbins, eta and frequency range (explicit bseg values carried over, the
    /// defaults re-derived).
    SearchConfig updated(SizeType nbins, double eta, double f_min, double f_max) const;

private:
    SearchConfigArgs m_args;
    SizeType m_bseg_brute{};
    SizeType m_bseg_ffa{};
    double m_tseg_brute{};
    double m_tseg_ffa{};
    SizeType m_niters_ffa{};
    std::vector<SizeType> m_widths;

    void validate() const;
};

}  // namespace loki_amd
