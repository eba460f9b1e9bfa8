// Region planner of a frequency sweep (regions.cpp:31-420).
#include "regions.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <utility>

namespace loki_amd {

namespace {
constexpr float kGiBf = 1073741824.0F;
}  // namespace

std::vector<FFARegion> generate_ffa_regions(double p_min, double p_max, double tsamp,
                                            SizeType nbins_min, double eta_min,
                                            double octave_scale, SizeType nbins_max) {
    LOKI_REQUIRE(p_min > 0.0, "p_min must be positive.");
    LOKI_REQUIRE(p_max > p_min, "p_max must be > p_min.");
    LOKI_REQUIRE(tsamp > 0.0, "tsamp must be positive.");
    LOKI_REQUIRE(nbins_min >= 2, "nbins_min must be >= 2.");
    LOKI_REQUIRE(eta_min > 0.0, "eta_min must be positive.");
    LOKI_REQUIRE(octave_scale >= 1.0, "octave_scale must be >= 1.0.");
    if (nbins_max == 0) {
        nbins_max = static_cast<SizeType>(-1);
    }
    LOKI_REQUIRE(nbins_max >= nbins_min, "nbins_max must be >= nbins_min.");
    std::vector<FFARegion> out;
    // one folding bin keeps its physical width p_min / nbins across bands, and
    // eta scales with nbins so the duty-cycle resolution rho stays fixed
    SizeType nb = std::min(nbins_min, static_cast<SizeType>(p_min / tsamp));
    const double rho       = eta_min / static_cast<double>(nb);
    const double bin_width = p_min / static_cast<double>(nb);
    double p_lo            = p_min;
    while (p_lo < p_max) {
        if (nb >= nbins_max) {
            out.push_back({1.0 / p_max, 1.0 / p_lo, nbins_max,
                           std::round(rho * static_cast<double>(nbins_max))});
            break;
        }
        const double p_hi = std::min(p_lo * octave_scale, p_max);
        out.push_back({1.0 / p_hi, 1.0 / p_lo, nb, rho * static_cast<double>(nb)});
        p_lo = p_hi;
        nb   = static_cast<SizeType>(p_lo / bin_width);
    }
    return out;
}

FFARegionStats::FFARegionStats(SizeType max_buffer_size, SizeType max_coord_size,
                               SizeType max_ncoords, SizeType max_ffa_levels, SizeType n_widths,
                               SizeType n_params, SizeType n_samps,
                               SizeType max_passing_candidates, bool use_fourier, bool use_gpu)
    : m_max_buffer_size(max_buffer_size), m_max_coord_size(max_coord_size),
      m_max_ncoords(max_ncoords), m_max_ffa_levels(max_ffa_levels), m_n_widths(n_widths),
      m_n_params(n_params), m_n_samps(n_samps), m_max_passing_candidates(max_passing_candidates),
      m_use_fourier(use_fourier), m_use_gpu(use_gpu) {}

SizeType FFARegionStats::max_buffer_size_time() const noexcept {
    return m_use_fourier ? 2 * m_max_buffer_size : m_max_buffer_size;
}
SizeType FFARegionStats::max_scores_size() const noexcept {
    return std::max(m_max_ncoords * m_n_widths, m_max_passing_candidates);
}
SizeType FFARegionStats::write_param_sets_size() const noexcept {
    return kFFAFreqSweepWriteBatchSize * (m_n_params + 1);
}
// The byte counts become GiB in single precision, as the reference computes
// them (regions.cpp:110-152): the planner's fit decisions then agree with it
// at the boundary.
float FFARegionStats::buffer_memory_usage() const noexcept {
    const SizeType elem = m_use_fourier ? 8 : 4;  // complex64 or float
    return static_cast<float>(2 * m_max_buffer_size * elem) / kGiBf;
}
float FFARegionStats::coord_memory_usage() const noexcept {
    const SizeType rec = m_n_params == 1 ? sizeof(FFACoordFreq) : sizeof(FFACoord);
    return static_cast<float>(m_max_coord_size * rec) / kGiBf;
}
float FFARegionStats::extra_memory_usage() const noexcept {
    return static_cast<float>(write_param_sets_size() * sizeof(double) +
                              max_scores_size() * sizeof(uint32_t) +
                              max_scores_size() * sizeof(float)) /
           kGiBf;
}
float FFARegionStats::cpu_memory_usage() const noexcept {
    return buffer_memory_usage() + coord_memory_usage() + extra_memory_usage();
}
float FFARegionStats::device_memory_usage() const noexcept {
    // both series, the candidate scores and indices, the fold buffers and tables
    const float extra = static_cast<float>((2 * m_n_samps + max_scores_size()) * sizeof(float) +
                                           max_scores_size() * sizeof(uint32_t)) /
                        kGiBf;
    return extra + buffer_memory_usage() + coord_memory_usage();
}
float FFARegionStats::freq_sweep_memory_usage() const noexcept {
    return m_use_gpu ? device_memory_usage() : cpu_memory_usage();
}

FFARegionPlanner::FFARegionPlanner(const SearchConfig& cfg, FoldKind kind, bool use_gpu)
    : m_base(cfg), m_kind(kind), m_use_gpu(use_gpu) {
    const double p_min  = 1.0 / m_base.get_f_max();
    const double p_max  = 1.0 / m_base.get_f_min();
    const auto regions  = generate_ffa_regions(p_min, p_max, m_base.get_tsamp(),
                                               m_base.get_nbins(), m_base.get_eta(),
                                               m_base.get_octave_scale(), m_base.get_nbins_max());
    const double drift = max_drift();
    LOKI_REQUIRE(drift >= 0.0 && drift < 1.0,
                 "FFARegionPlanner: max_drift must be in [0, 1); the parameter ranges are too "
                 "large for the observation duration.");
    SizeType mbuf = 0, mcoord = 0, mnc = 0, mlv = 0;
    for (const auto& r : regions) {
        subdivide(r, drift, mbuf, mcoord, mnc, mlv);
    }
    m_stats = FFARegionStats(mbuf, mcoord, mnc, mlv, m_base.get_scoring_widths().size(),
                             m_base.get_nparams(), m_base.get_nsamps(),
                             m_base.get_max_passing_candidates(), m_base.get_use_fourier(),
                             m_use_gpu);
}

// regions.cpp:181-207: largest fractional Doppler shift over half the
// observation from the accel (and jerk) ranges.
double FFARegionPlanner::max_drift() const {
    const SizeType np = m_base.get_nparams();
    if (np <= 1) {
        return 0.0;
    }
    const auto lim    = m_base.get_param_limits();
    const double half = m_base.get_tobs() / 2.0;
    auto amax = [](const ParamLimit& l) { return std::max(std::abs(l.min), std::abs(l.max)); };
    constexpr double kC = 299792458.0;
    if (np == 2) {
        return amax(lim[0]) * half / kC;
    }
    LOKI_REQUIRE(np == 3, "Unsupported number of parameters for drift calculation");
    return (amax(lim[1]) * half + amax(lim[0]) * half * half / 2.0) / kC;
}

// regions.cpp:247-420: cover [f_start, f_end] from the top down with the
// widest chunks whose drift-widened FFA fits the memory limit (bisection on
// the width), merging a thin low-frequency remainder into its neighbour.
void FFARegionPlanner::subdivide(const FFARegion& r, double drift, SizeType& mbuf,
                                 SizeType& mcoord, SizeType& mnc, SizeType& mlv) {
    if (r.f_end <= r.f_start) {
        return;
    }
    LOKI_REQUIRE(r.f_start * (1.0 - drift) > 0.0,
                 "FFARegionPlanner: drift expansion would push f_start to a non-positive "
                 "frequency. Reduce parameter ranges or raise f_min.");
    const double limit_gb = m_base.get_max_process_memory_gb() - 0.5;  // headroom
    LOKI_REQUIRE(limit_gb > 0.0,
                 "FFARegionPlanner: max_process_memory_gb must exceed the 0.5 GB margin.");
    const double tol =
        std::max(1.0e-2, 1.0e-4 * (r.f_end - r.f_start));  // boundary resolution (Hz)

    struct Eval {
        SearchConfig cfg;
        SizeType buf, coord, nc, lv;
        double own_gb, alloc_gb;
    };
    auto stats_of = [&](SizeType b, SizeType c, SizeType n, SizeType l) {
        return FFARegionStats(b, c, n, l, m_base.get_scoring_widths().size(), m_base.get_nparams(),
                              m_base.get_nsamps(), m_base.get_max_passing_candidates(),
                              m_base.get_use_fourier(), m_use_gpu);
    };
    auto evaluate = [&](double lo, double hi) {
        LOKI_REQUIRE(hi > lo, "FFARegionPlanner: zero/negative-width chunk.");
        const double a = lo * (1.0 - drift);
        const double b = hi * (1.0 + drift);
        LOKI_REQUIRE(a > 0.0 && b > a, "FFARegionPlanner: invalid drift-expanded range.");
        SearchConfig cfg = m_base.updated(r.nbins, r.eta, a, b);
        const FFAPlan plan(cfg, m_kind);
        const SizeType buf = plan.buffer_size(), coord = plan.coord_size();
        const SizeType nc = plan.ncoords().back(), lv = plan.n_levels();
        const double own   = stats_of(buf, coord, nc, lv).freq_sweep_memory_usage();
        // the sweep's workspace holds the maxima over all chunks
        const double alloc = stats_of(std::max(mbuf, buf), std::max(mcoord, coord),
                                      std::max(mnc, nc), std::max(mlv, lv))
                                 .freq_sweep_memory_usage();
        return Eval{std::move(cfg), buf, coord, nc, lv, own, alloc};
    };
    auto fits = [&](const Eval& e) { return e.alloc_gb <= limit_gb; };

    double hi = r.f_end;
    while (hi > r.f_start) {
        double lo = r.f_start;
        Eval best = evaluate(r.f_start, hi);
        if (!fits(best)) {
            const double min_w = std::min(hi - r.f_start, 1.0e-2);
            best               = evaluate(hi - min_w, hi);
            LOKI_REQUIRE(fits(best),
                         "FFARegionPlanner: cannot fit the minimum viable chunk; increase "
                         "max_process_memory_gb or reduce the parameter ranges.");
            double w_ok = min_w, w_bad = hi - r.f_start;
            for (int step = 0; step < 50 && (w_bad - w_ok) > tol; ++step) {
                const double w = std::midpoint(w_ok, w_bad);
                Eval e         = evaluate(hi - w, hi);
                if (fits(e)) {
                    w_ok = w;
                    best = std::move(e);
                } else {
                    w_bad = w;
                }
            }
            lo = hi - w_ok;
            // a remainder of at most 4 tolerances is absorbed when that still fits
            const double rest = lo - r.f_start;
            if (rest > 0.0 && rest <= 4.0 * tol) {
                Eval merged = evaluate(r.f_start, hi);
                if (fits(merged)) {
                    lo   = r.f_start;
                    best = std::move(merged);
                }
            }
        }
        LOKI_REQUIRE(lo < hi, "FFARegionPlanner: no progress (planner bug).");
        const double a = lo * (1.0 - drift), b = hi * (1.0 + drift);
        m_cfgs.push_back(best.cfg);
        m_chunks.push_back({lo, hi, a, b, hi - lo, b - a, best.own_gb,
                            ((b - a) - (hi - lo)) / (b - a)});
        mbuf   = std::max(mbuf, best.buf);
        mcoord = std::max(mcoord, best.coord);
        mnc    = std::max(mnc, best.nc);
        mlv    = std::max(mlv, best.lv);
        hi     = lo;
    }
}

}  // namespace loki_amd
