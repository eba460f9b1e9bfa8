// PulsarSearchConfig restatement (lib/configs.cpp:18-303) and the host grid
// helpers (lib/psr_utils.cpp:14-45, lib/loki/psr_utils.hpp:235-270,
// lib/score.cpp:390-405).
#include "config.hpp"

#include <algorithm>
#include <array>
#include <cmath>

#include "grid.hpp"

namespace loki_amd {

namespace {
bool is_power_of_2(SizeType n) { return n > 0 && (n & (n - 1)) == 0; }
constexpr std::array<const char*, 5> kParamNames = {"crackle", "snap", "jerk", "accel",
                                                    "freq"};
}  // namespace

std::vector<double> poly_taylor_step_f(SizeType nparams, double tobs, SizeType fold_bins,
                                       double tol_bins, double t_ref) {
    const double dphi = tol_bins / static_cast<double>(fold_bins);
    const double dt   = tobs - t_ref;
    std::vector<double> out(nparams);
    for (SizeType i = 0; i < nparams; ++i) {
        const double orth = std::pow(2.0, static_cast<double>(i));
        // math.hpp:33 factorial(double) == tgamma(n + 1)
        const double fact = std::tgamma(static_cast<double>(i + 1) + 1.0);
        out[nparams - 1 - i] = dphi * orth * fact / std::pow(dt, static_cast<double>(i + 1));
    }
    return out;
}

std::vector<double> poly_taylor_step_d_f(SizeType nparams, double tobs, SizeType fold_bins,
                                         double tol_bins, double f_max, double t_ref) {
    const auto df = poly_taylor_step_f(nparams, tobs, fold_bins, tol_bins, t_ref);
    std::vector<double> out(nparams);
    for (SizeType i = 0; i + 1 < nparams; ++i) {
        out[i] = df[i] * grid::kCval / f_max;
    }
    out[nparams - 1] = df[nparams - 1];
    return out;
}

SizeType range_param_count(double vmin, double vmax, double dv) {
    LOKI_REQUIRE(vmax > vmin, "vmax must be greater than vmin");
    LOKI_REQUIRE(dv > 0, "dv must be positive");
    if (dv >= (vmax - vmin)) {
        return 1;
    }
    return static_cast<SizeType>(std::ceil((vmax - vmin) / dv));
}

std::vector<double> range_param(double vmin, double vmax, double dv) {
    LOKI_REQUIRE(vmax > vmin, "vmax must be greater than vmin");
    LOKI_REQUIRE(dv > 0, "dv must be positive");
    if (dv >= (vmax - vmin)) {
        return {(vmax + vmin) / 2.0};
    }
    const auto npoints     = static_cast<SizeType>(std::ceil((vmax - vmin) / dv));
    const double dv_actual = (vmax - vmin) / static_cast<double>(npoints);
    const double start     = vmin + (0.5 * dv_actual);
    std::vector<double> out(npoints);
    for (SizeType i = 0; i < npoints; ++i) {
        out[i] = start + (dv_actual * static_cast<double>(i));
    }
    return out;
}

std::vector<SizeType> generate_box_width_trials(SizeType fold_bins, double ducy_max,
                                                double wtsp) {
    const auto wmax =
        static_cast<SizeType>(std::max(1.0, ducy_max * static_cast<double>(fold_bins)));
    std::vector<SizeType> widths = {1};
    while (widths.back() < wmax) {
        const auto next = std::max(static_cast<SizeType>(widths.back() + 1),
                                   static_cast<SizeType>(wtsp * static_cast<double>(widths.back())));
        if (next > wmax) {
            break;
        }
        widths.push_back(next);
    }
    return widths;
}

SearchConfig::SearchConfig(SearchConfigArgs args) : m_args(std::move(args)) {
    LOKI_REQUIRE(!m_args.param_limits.empty(), "coord_limits must be non-empty");
    LOKI_REQUIRE(m_args.param_limits.size() <= kParamNames.size(),
                 "at most 5 search parameters are supported");
    LOKI_REQUIRE(m_args.nsamps > 0, "nsamps must be positive");
    LOKI_REQUIRE(is_power_of_2(m_args.nsamps), "nsamps must be a power of 2");
    LOKI_REQUIRE(m_args.tsamp > 0, "tsamp must be positive");
    LOKI_REQUIRE(m_args.nbins > 0, "nbins must be positive");
    const SizeType nparams = m_args.param_limits.size();
    if (m_args.bseg_brute.has_value()) {
        m_bseg_brute = *m_args.bseg_brute;
    } else {
        // configs.cpp:295-300
        const SizeType init_levels = (nparams == 1) ? 1 : 2;
        const double lv = std::log2(static_cast<double>(m_args.nsamps) * m_args.tsamp *
                                    get_f_min());
        LOKI_REQUIRE(std::isfinite(lv) && lv >= static_cast<double>(init_levels),
                     "cannot derive a default bseg_brute (nsamps*tsamp*f_min too small)");
        const auto levels = static_cast<SizeType>(lv);
        m_bseg_brute      = m_args.nsamps / (SizeType{1} << (levels - init_levels));
    }
    m_bseg_ffa = m_args.bseg_ffa.value_or(m_args.nsamps);
    m_args.nthreads = std::max(m_args.nthreads, 1);
    validate();
    m_tseg_brute = static_cast<double>(m_bseg_brute) * m_args.tsamp;
    m_tseg_ffa   = static_cast<double>(m_bseg_ffa) * m_args.tsamp;
    m_niters_ffa = static_cast<SizeType>(
        std::log2(static_cast<double>(m_bseg_ffa) / static_cast<double>(m_bseg_brute)));
    m_widths = generate_box_width_trials(m_args.nbins, m_args.ducy_max, m_args.wtsp);
}

void SearchConfig::validate() const {
    // configs.cpp:262-293
    LOKI_REQUIRE(is_power_of_2(m_args.nsamps), "nsamps must be a power of 2");
    LOKI_REQUIRE(m_args.tsamp > 0, "tsamp must be positive");
    LOKI_REQUIRE(m_args.eta > 0, "eta (tolerance bins) must be positive");
    LOKI_REQUIRE(m_args.max_process_memory_gb > 0, "max_process_memory_gb must be positive");
    LOKI_REQUIRE(m_args.nbins_max >= m_args.nbins,
                 "nbins_max must be greater than or equal to nbins");
    LOKI_REQUIRE(is_power_of_2(m_bseg_brute), "bseg_brute must be a power of 2");
    LOKI_REQUIRE(is_power_of_2(m_bseg_ffa), "bseg_ffa must be a power of 2");
    LOKI_REQUIRE(m_bseg_brute < m_args.nsamps, "bseg_brute must be less than nsamps");
    LOKI_REQUIRE(m_bseg_ffa <= m_args.nsamps, "bseg_ffa must be less than or equal to nsamps");
    LOKI_REQUIRE(m_bseg_ffa >= m_bseg_brute,
                 "bseg_ffa must be greater than or equal to bseg_brute");
    for (const auto& lim : m_args.param_limits) {
        LOKI_REQUIRE(lim.max >= lim.min, "param_limits must be increasing");
    }
}

std::vector<std::string> SearchConfig::get_param_names() const {
    const SizeType n = get_nparams();
    std::vector<std::string> names;
    for (SizeType i = kParamNames.size() - n; i < kParamNames.size(); ++i) {
        names.emplace_back(kParamNames[i]);
    }
    return names;
}

std::vector<double> SearchConfig::get_dparams_f(double tseg_cur) const {
    const double t_ref = (get_nparams() == 1) ? 0.0 : tseg_cur / 2.0;
    return poly_taylor_step_f(get_nparams(), tseg_cur, m_args.nbins, m_args.eta, t_ref);
}

std::vector<double> SearchConfig::get_dparams(double tseg_cur) const {
    const double t_ref = (get_nparams() == 1) ? 0.0 : tseg_cur / 2.0;
    return poly_taylor_step_d_f(get_nparams(), tseg_cur, m_args.nbins, m_args.eta,
                                get_f_max(), t_ref);
}

std::vector<SizeType> SearchConfig::get_param_grid_count(double tseg_cur) const {
    const auto dp = get_dparams(tseg_cur);
    std::vector<SizeType> count(get_nparams());
    for (SizeType i = 0; i < get_nparams(); ++i) {
        count[i] = range_param_count(m_args.param_limits[i].min, m_args.param_limits[i].max,
                                     dp[i]);
    }
    return count;
}

std::vector<double> SearchConfig::get_dparams_actual(double tseg_cur) const {
    const auto cnt = get_param_grid_count(tseg_cur);
    std::vector<double> out(get_nparams());
    for (SizeType i = 0; i < get_nparams(); ++i) {
        out[i] = (m_args.param_limits[i].max - m_args.param_limits[i].min) /
                 static_cast<double>(cnt[i]);
    }
    return out;
}

SearchConfig SearchConfig::updated(SizeType nbins, double eta, double f_min,
                                   double f_max) const {
    SearchConfigArgs a  = m_args;
    a.nbins             = nbins;
    a.eta               = eta;
    a.param_limits.back() = ParamLimit{f_min, f_max};
    return SearchConfig(std::move(a));
}

}  // namespace loki_amd
