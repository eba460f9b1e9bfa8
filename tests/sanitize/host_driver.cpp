// Host-side driver for the AddressSanitizer / UndefinedBehaviorSanitizer build
// of the library's host C++ (config.cpp, plan.cpp, regions.cpp).
//
// tests/test_sanitize.py compiles this file together with those three sources
// under -fsanitize=address,undefined and runs it.  Each input line names one
// case; the driver builds the configuration, its time or Fourier plan (and
// for "planner" lines the region planner) and prints one JSON object per
// case with the figures the test compares against the oracle's restatement.
// Invalid configurations print {"error": "..."}: the error paths run under
// the sanitizers too.
//
// Input lines (whitespace separated):
//   plan <fourier 0|1> <nsamps> <tsamp> <nbins> <eta> <nparams> <lo hi>...
//   planner <fourier 0|1> <mem_gb> <nsamps> <tsamp> <nbins> <eta> <nparams> <lo hi>...
//   regions <p_min> <p_max> <tsamp> <nbins_min> <eta_min> <octave_scale> <nbins_max>
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "config.hpp"
#include "plan.hpp"
#include "regions.hpp"

using namespace loki_amd;

namespace {

void put_sizes(std::ostream& os, const std::vector<SizeType>& v) {
    os << '[';
    for (size_t i = 0; i < v.size(); ++i) {
        os << (i ? "," : "") << v[i];
    }
    os << ']';
}

void put_doubles(std::ostream& os, const std::vector<double>& v) {
    os << '[';
    char buf[40];
    for (size_t i = 0; i < v.size(); ++i) {
        std::snprintf(buf, sizeof buf, "%.17g", v[i]);
        os << (i ? "," : "") << buf;
    }
    os << ']';
}

SearchConfigArgs read_args(std::istringstream& in, bool fourier) {
    SearchConfigArgs a;
    SizeType np = 0;
    in >> a.nsamps >> a.tsamp >> a.nbins >> a.eta >> np;
    for (SizeType k = 0; k < np; ++k) {
        ParamLimit p{};
        in >> p.min >> p.max;
        a.param_limits.push_back(p);
    }
    a.use_fourier = fourier;
    return a;
}

void plan_case(std::istringstream& in) {
    int fourier = 0;
    in >> fourier;
    const SearchConfig cfg(read_args(in, fourier != 0));
    const FFAPlan plan(cfg, fourier ? FoldKind::kFourier : FoldKind::kTime);
    std::ostringstream os;
    os << "{\"n_levels\":" << plan.n_levels() << ",\"segment_lens\":";
    put_sizes(os, plan.segment_lens());
    os << ",\"nsegments\":";
    put_sizes(os, plan.nsegments());
    os << ",\"ncoords\":";
    put_sizes(os, plan.ncoords());
    os << ",\"fold_shapes\":[";
    for (size_t l = 0; l < plan.fold_shapes().size(); ++l) {
        os << (l ? "," : "");
        put_sizes(os, plan.fold_shapes()[l]);
    }
    os << "],\"dparams\":[";
    for (size_t l = 0; l < plan.dparams().size(); ++l) {
        os << (l ? "," : "");
        put_doubles(os, plan.dparams()[l]);
    }
    os << "],\"widths\":";
    put_sizes(os, cfg.get_scoring_widths());
    os << ",\"buffer_size\":" << plan.buffer_size() << ",\"fold_size\":" << plan.fold_size()
       << ",\"coord_size\":" << plan.coord_size() << ",\"bseg_brute\":" << cfg.get_bseg_brute()
       << ",\"freqs\":";
    // every level's frequency grid (compute_param_grid walks the widest arrays)
    std::vector<double> last;
    for (SizeType l = 0; l < plan.n_levels(); ++l) {
        last = plan.compute_param_grid(l).back();
    }
    put_doubles(os, last);
    os << "}";
    std::cout << os.str() << "\n";
}

void planner_case(std::istringstream& in) {
    int fourier = 0;
    double mem  = 0;
    in >> fourier >> mem;
    SearchConfigArgs a             = read_args(in, fourier != 0);
    a.max_process_memory_gb        = mem;
    a.max_passing_candidates       = 1U << 14U;
    const SearchConfig cfg(a);
    const FFARegionPlanner planner(cfg, fourier ? FoldKind::kFourier : FoldKind::kTime, true);
    std::ostringstream os;
    os << "{\"nregions\":" << planner.nregions() << ",\"nbins\":[";
    for (size_t i = 0; i < planner.cfgs().size(); ++i) {
        os << (i ? "," : "") << planner.cfgs()[i].get_nbins();
    }
    os << "],\"f_lo\":";
    std::vector<double> lo, hi;
    for (const auto& c : planner.cfgs()) {
        lo.push_back(c.get_f_min());
        hi.push_back(c.get_f_max());
    }
    put_doubles(os, lo);
    os << ",\"f_hi\":";
    put_doubles(os, hi);
    const auto& st = planner.stats();
    os << ",\"max_buffer_size\":" << st.max_buffer_size()
       << ",\"max_coord_size\":" << st.max_coord_size() << ",\"max_ncoords\":" << st.max_ncoords()
       << ",\"max_ffa_levels\":" << st.max_ffa_levels() << "}";
    std::cout << os.str() << "\n";
}

void regions_case(std::istringstream& in) {
    double p_min = 0, p_max = 0, tsamp = 0, eta = 0, scale = 0;
    SizeType nbins_min = 0, nbins_max = 0;
    in >> p_min >> p_max >> tsamp >> nbins_min >> eta >> scale >> nbins_max;
    const auto rs = generate_ffa_regions(p_min, p_max, tsamp, nbins_min, eta, scale, nbins_max);
    std::ostringstream os;
    std::vector<double> fs, fe;
    std::vector<SizeType> nb;
    for (const auto& r : rs) {
        fs.push_back(r.f_start);
        fe.push_back(r.f_end);
        nb.push_back(r.nbins);
    }
    os << "{\"f_start\":";
    put_doubles(os, fs);
    os << ",\"f_end\":";
    put_doubles(os, fe);
    os << ",\"nbins\":";
    put_sizes(os, nb);
    os << "}";
    std::cout << os.str() << "\n";
}

}  // namespace

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        if (line.empty()) {
            continue;
        }
        std::istringstream in(line);
        std::string kind;
        in >> kind;
        try {
            if (kind == "plan") {
                plan_case(in);
            } else if (kind == "planner") {
                planner_case(in);
            } else if (kind == "regions") {
                regions_case(in);
            } else {
                std::cout << "{\"error\":\"unknown case\"}\n";
            }
        } catch (const std::exception& e) {
            std::string msg = e.what();
            for (auto& ch : msg) {
                if (ch == '"' || ch == '\\' || ch < 32) {
                    ch = ' ';
                }
            }
            std::cout << "{\"error\":\"" << msg << "\"}\n";
        }
    }
    return 0;
}
