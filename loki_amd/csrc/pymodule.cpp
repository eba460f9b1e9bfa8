// Python bindings over the C-ABI (include/loki_amd.h), mirroring the names,
// argument names/defaults and return types of loki's pybind modules:
//   libloki:   configs.PulsarSearchConfig, plans.FFAPlan{Base,Time,Fourier},
//              plans.generate_ffa_regions / FFARegionStats / FFARegionPlanner{Time,Fourier},
//              scores.generate_box_width_trials   (src/loki_python.cpp:71-520)
//   libculoki: ffa.FFACUDA / FFACOMPLEXCUDA / compute_ffa_time_cuda /
//              compute_ffa_fourier_cuda / compute_ffa_fourier_return_to_time_cuda /
//              compute_ffa_scores_cuda, FFAFreqSweepCUDA,
//              fold.compute_brute_fold_{time,fourier}_cuda,
//              scores.snr_boxcar_*
//              (src/loki_python_cuda.cpp:18-203, loki_templates_cuda.cuh:21-51)
// C-ABI failures raise RuntimeError with the library's message, the same
// Python exception type pybind11 gives for the reference's C++ exceptions.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "../../include/loki_amd.h"

namespace py = pybind11;

template <typename T>
using PyArrayT = py::array_t<T, py::array::c_style | py::array::forcecast>;

namespace {

void check(int rc) {
    if (rc != LOKI_OK) {
        throw std::runtime_error(loki_last_error());
    }
}

template <typename T> py::array_t<T> to_array(const std::vector<T>& v) {
    py::array_t<T> a(static_cast<py::ssize_t>(v.size()));
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

const std::vector<std::string> kNames = {"crackle", "snap", "jerk", "accel", "freq"};

class Config {
public:
    Config(size_t nsamps, double tsamp, size_t nbins, double eta,
           const PyArrayT<double>& param_limits, double ducy_max, double wtsp, bool use_fourier,
           int nthreads, double max_process_memory_gb, double octave_scale, size_t nbins_max,
           size_t nbins_min_lossy_bf, std::optional<size_t> bseg_brute,
           std::optional<size_t> bseg_ffa, double snr_min, size_t max_passing_candidates,
           size_t prune_poly_order, double p_orb_min, double m_c_max, double m_p_min,
           double minimum_snap_cells, bool use_conservative_tile) {
        if (param_limits.ndim() != 2 || param_limits.shape(1) != 2) {
            throw std::invalid_argument(
                "param_limits must be a 2D NumPy array with shape (n_params, 2)");
        }
        loki_config_params p{};
        p.nsamps                 = nsamps;
        p.tsamp                  = tsamp;
        p.nbins                  = nbins;
        p.eta                    = eta;
        p.param_limits           = param_limits.data();
        p.nparams                = static_cast<size_t>(param_limits.shape(0));
        p.ducy_max               = ducy_max;
        p.wtsp                   = wtsp;
        p.use_fourier            = use_fourier ? 1 : 0;
        p.nthreads               = nthreads;
        p.max_process_memory_gb  = max_process_memory_gb;
        p.octave_scale           = octave_scale;
        p.nbins_max              = nbins_max;
        p.nbins_min_lossy_bf     = nbins_min_lossy_bf;
        p.bseg_brute             = bseg_brute.value_or(0);
        p.bseg_ffa               = bseg_ffa.value_or(0);
        p.snr_min                = snr_min;
        p.max_passing_candidates = max_passing_candidates;
        p.prune_poly_order       = prune_poly_order;
        p.p_orb_min              = p_orb_min;
        p.m_c_max                = m_c_max;
        p.m_p_min                = m_p_min;
        p.minimum_snap_cells     = minimum_snap_cells;
        p.use_conservative_tile  = use_conservative_tile ? 1 : 0;
        loki_config_t h          = nullptr;
        check(loki_config_create(&p, &h));
        m_h.reset(h, loki_config_destroy);
        check(loki_config_info_get(h, &m_info));
        check(loki_config_params_get(h, &m_p));
    }
    /// Takes ownership of a handle made by the library (e.g. a planned chunk).
    explicit Config(loki_config_t h) {
        m_h.reset(h, loki_config_destroy);
        check(loki_config_info_get(h, &m_info));
        check(loki_config_params_get(h, &m_p));
    }
    loki_config_t handle() const { return m_h.get(); }
    const loki_config_info& info() const { return m_info; }
    const loki_config_params& params() const { return m_p; }
    py::array_t<double> param_limits() const {
        py::array_t<double> a({static_cast<py::ssize_t>(m_p.nparams), py::ssize_t{2}});
        std::copy(m_p.param_limits, m_p.param_limits + 2 * m_p.nparams, a.mutable_data());
        return a;
    }
    std::vector<std::string> param_names() const {
        return {kNames.end() - static_cast<long>(m_p.nparams), kNames.end()};
    }
    py::array_t<size_t> score_widths() const {
        size_t n = 0;
        check(loki_config_scoring_widths(handle(), nullptr, 0, &n));
        std::vector<size_t> w(n);
        check(loki_config_scoring_widths(handle(), w.data(), n, &n));
        return to_array(w);
    }
    std::vector<double> dparams(double tseg, int kind) const {
        std::vector<double> v(m_p.nparams);
        check(loki_config_dparams(handle(), tseg, kind, v.data()));
        return v;
    }

private:
    std::shared_ptr<loki_config_s> m_h;
    loki_config_info m_info{};
    loki_config_params m_p{};
};

class Plan {
public:
    Plan(const Config& cfg, int fold_type) : m_cfg(cfg), m_fold_type(fold_type) {
        loki_plan_t h = nullptr;
        check(loki_plan_create(cfg.handle(), fold_type, &h));
        m_h.reset(h, loki_plan_destroy);
        check(loki_plan_info_get(h, &m_info));
        m_levels.resize(m_info.n_levels);
        for (size_t l = 0; l < m_info.n_levels; ++l) {
            check(loki_plan_level_get(h, l, &m_levels[l]));
        }
    }
    const loki_plan_info& info() const { return m_info; }
    const Config& config() const { return m_cfg; }
    int fold_type() const { return m_fold_type; }
    template <typename F> py::array level_array(F get) const {
        using T = decltype(get(m_levels[0]));
        std::vector<T> v;
        for (const auto& lv : m_levels) {
            v.push_back(get(lv));
        }
        return to_array(v);
    }
    py::list per_level(bool counts, bool strides, bool dp, bool dpa) const {
        py::list out;
        for (const auto& lv : m_levels) {
            if (counts || strides) {
                std::vector<size_t> v(lv.param_counts, lv.param_counts + m_info.n_params);
                if (strides) {
                    v.assign(lv.param_cart_strides, lv.param_cart_strides + m_info.n_params);
                }
                out.append(to_array(v));
            } else {
                const double* src = dp ? lv.dparams : lv.dparams_actual;
                (void)dpa;
                out.append(to_array(std::vector<double>(src, src + m_info.n_params)));
            }
        }
        return out;
    }
    py::list fold_shapes(bool time) const {
        py::list out;
        for (const auto& lv : m_levels) {
            const size_t* s = time ? lv.fold_shape_time : lv.fold_shape;
            out.append(to_array(std::vector<size_t>(s, s + lv.fold_ndim)));
        }
        return out;
    }
    std::vector<double> param_grid(size_t level, size_t param) const {
        size_t n = 0;
        check(loki_plan_param_grid(m_h.get(), level, param, nullptr, 0, &n));
        std::vector<double> v(n);
        check(loki_plan_param_grid(m_h.get(), level, param, v.data(), n, &n));
        return v;
    }
    py::list compute_param_grid(size_t level) const {
        py::list out;
        for (size_t p = 0; p < m_info.n_params; ++p) {
            out.append(to_array(param_grid(level, p)));
        }
        return out;
    }
    py::dict params_dict() const {
        py::dict d;
        const auto names = m_cfg.param_names();
        for (size_t p = 0; p < m_info.n_params; ++p) {
            d[py::str(names[p])] = to_array(param_grid(m_info.n_levels - 1, p));
        }
        return d;
    }
    py::list resolve_coordinates(bool freq_only, int device_id) const {
        if (freq_only != (m_info.n_params == 1)) {
            throw std::runtime_error(freq_only
                                         ? "resolve_coordinates_freq() only supports nparams=1"
                                         : "resolve_coordinates only supports nparams>=2. For "
                                           "frequency coordinates, use "
                                           "resolve_coordinates_freq() instead.");
        }
        const size_t rec = freq_only ? 8 : 16;
        std::vector<uint8_t> buf(m_info.coord_size * rec);
        check(loki_plan_resolve_coordinates_host(m_h.get(), buf.data(), buf.size(), device_id));
        py::list out;
        py::list fields;
        if (freq_only) {
            fields.append(py::make_tuple("idx", "<u4"));
            fields.append(py::make_tuple("shift", "<f4"));
        } else {
            fields.append(py::make_tuple("i_tail", "<u4"));
            fields.append(py::make_tuple("shift_tail", "<f4"));
            fields.append(py::make_tuple("i_head", "<u4"));
            fields.append(py::make_tuple("shift_head", "<f4"));
        }
        py::dtype dt = py::module_::import("numpy").attr("dtype")(fields).cast<py::dtype>();
        for (const auto& lv : m_levels) {
            py::array a(dt, std::vector<py::ssize_t>{static_cast<py::ssize_t>(lv.ncoords)});
            std::memcpy(a.mutable_data(), buf.data() + lv.ncoords_offset * rec, lv.ncoords * rec);
            out.append(a);
        }
        return out;
    }

private:
    Config m_cfg;
    int m_fold_type;
    std::shared_ptr<loki_plan_s> m_h;
    loki_plan_info m_info{};
    std::vector<loki_plan_level> m_levels;
};

struct PlanTime : Plan {
    explicit PlanTime(const Config& c) : Plan(c, LOKI_FOLD_TIME) {}
};
struct PlanFourier : Plan {
    explicit PlanFourier(const Config& c) : Plan(c, LOKI_FOLD_FOURIER) {}
};

/// workspace.hpp:169-206 FFAWorkspaceCUDA: device buffers FFAs can share.
class Workspace {
public:
    Workspace(int device_id, size_t buffer_bytes, size_t coord_bytes) : m_device(device_id) {
        loki_ffa_workspace_t h = nullptr;
        check(loki_ffa_workspace_create(device_id, buffer_bytes, coord_bytes, &h));
        m_h.reset(h, loki_ffa_workspace_destroy);
    }
    loki_ffa_workspace_t handle() const { return m_h.get(); }
    int device() const { return m_device; }

private:
    int m_device;
    std::shared_ptr<loki_ffa_workspace_s> m_h;
};

class FFA {
public:
    FFA(const Config& cfg, int device_id, int fold_type, const Workspace* ws = nullptr)
        : m_cfg(cfg), m_device(device_id), m_fold_type(fold_type), m_plan(cfg, fold_type) {
        loki_ffa_t h = nullptr;
        if (ws != nullptr) {
            check(loki_ffa_create_with_workspace(cfg.handle(), fold_type, ws->handle(),
                                                 device_id, &h));
        } else {
            check(loki_ffa_create(cfg.handle(), fold_type, device_id, &h));
        }
        m_h.reset(h, loki_ffa_destroy);
    }
    const Plan& plan() const { return m_plan; }
    /// Time plan: float fold of buffer_size.  Fourier plan: float fold of
    /// 2*buffer_size, the final level returned to the time domain.
    void execute(const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v,
                 PyArrayT<float>& fold) {
        if (ts_v.size() != ts_e.size()) {
            throw std::runtime_error("FFA::execute: ts_v must have size nsamps");
        }
        if (m_fold_type == LOKI_FOLD_FOURIER) {
            check(loki_ffa_execute_time_host(m_h.get(), ts_e.data(), ts_v.data(),
                                             static_cast<size_t>(ts_e.size()),
                                             fold.mutable_data(),
                                             static_cast<size_t>(fold.size())));
            return;
        }
        check(loki_ffa_execute_host(m_h.get(), ts_e.data(), ts_v.data(),
                                    static_cast<size_t>(ts_e.size()), fold.mutable_data(),
                                    static_cast<size_t>(fold.size())));
    }
    /// Fourier plan: complex64 fold of buffer_size.
    void execute_complex(const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v,
                         PyArrayT<std::complex<float>>& fold) {
        if (ts_v.size() != ts_e.size()) {
            throw std::runtime_error("FFA::execute: ts_v must have size nsamps");
        }
        check(loki_ffa_execute_host(m_h.get(), ts_e.data(), ts_v.data(),
                                    static_cast<size_t>(ts_e.size()), fold.mutable_data(),
                                    static_cast<size_t>(fold.size())));
    }
    void execute_time_device(uintptr_t ts_e, uintptr_t ts_v, uintptr_t fold, size_t fold_len,
                             uintptr_t stream) {
        check(loki_ffa_execute_time_device(m_h.get(), reinterpret_cast<const float*>(ts_e),
                                           reinterpret_cast<const float*>(ts_v),
                                           reinterpret_cast<float*>(fold), fold_len,
                                           reinterpret_cast<void*>(stream)));
    }
    void execute_device(uintptr_t ts_e, uintptr_t ts_v, uintptr_t fold, size_t fold_len,
                        uintptr_t stream) {
        check(loki_ffa_execute_device(m_h.get(), reinterpret_cast<const float*>(ts_e),
                                      reinterpret_cast<const float*>(ts_v),
                                      reinterpret_cast<void*>(fold), fold_len,
                                      reinterpret_cast<void*>(stream)));
    }
    void scores_device(uintptr_t ts_e, uintptr_t ts_v, const PyArrayT<size_t>& widths,
                       uintptr_t scores, size_t scores_len, bool find_max, uintptr_t stream) {
        check(loki_ffa_scores_device(m_h.get(), reinterpret_cast<const float*>(ts_e),
                                     reinterpret_cast<const float*>(ts_v), widths.data(),
                                     static_cast<size_t>(widths.size()),
                                     reinterpret_cast<float*>(scores), scores_len,
                                     find_max ? 1 : 0, reinterpret_cast<void*>(stream)));
    }
    void scores_filter_device(uintptr_t ts_e, uintptr_t ts_v, const PyArrayT<size_t>& widths,
                              double threshold, bool find_max, uintptr_t indices,
                              uintptr_t scores, size_t capacity, uintptr_t count,
                              uintptr_t stream) {
        check(loki_ffa_scores_filter_device(
            m_h.get(), reinterpret_cast<const float*>(ts_e), reinterpret_cast<const float*>(ts_v),
            widths.data(), static_cast<size_t>(widths.size()), static_cast<float>(threshold),
            find_max ? 1 : 0, reinterpret_cast<uint32_t*>(indices),
            reinterpret_cast<float*>(scores), capacity, reinterpret_cast<uint32_t*>(count),
            reinterpret_cast<void*>(stream)));
    }
    void set_timing(bool on) { check(loki_ffa_set_timing(m_h.get(), on ? 1 : 0)); }
    void set_fused(bool on) { check(loki_ffa_set_fused(m_h.get(), on ? 1 : 0)); }
    void set_fuse_depth(int depth) { check(loki_ffa_set_fuse_depth(m_h.get(), depth)); }
    void set_two_level_merges(bool on) {
        check(loki_ffa_set_two_level_merges(m_h.get(), on ? 1 : 0));
    }
    void set_exact_norm(bool on) { check(loki_ffa_set_exact_norm(m_h.get(), on ? 1 : 0)); }
    void set_dft_precision(const std::string& p) {
        if (p != "f32" && p != "bf16x3") {
            throw std::invalid_argument("dft precision must be 'f32' or 'bf16x3'");
        }
        check(loki_ffa_set_dft_precision(m_h.get(), p == "bf16x3" ? 1 : 0));
    }
    void set_fold_storage(const std::string& st) {
        int k = -1;
        if (st == "f32" || st == "float32") {
            k = LOKI_STORAGE_F32;
        } else if (st == "f16" || st == "float16") {
            k = LOKI_STORAGE_F16;
        } else if (st == "bf16" || st == "bfloat16") {
            k = LOKI_STORAGE_BF16;
        } else if (st == "bfp16") {
            k = LOKI_STORAGE_BFP16;
        } else {
            throw std::runtime_error("fold_storage must be 'f32', 'f16', 'bf16' or 'bfp16'");
        }
        check(loki_ffa_set_fold_storage(m_h.get(), k));
    }
    py::list timings() {
        size_t n = 0;
        check(loki_ffa_timings(m_h.get(), nullptr, 0, &n));
        std::vector<loki_kernel_timing> t(n);
        check(loki_ffa_timings(m_h.get(), t.data(), n, &n));
        static const char* kinds[] = {"resolve", "brute_fold",  "merge",
                                      "score",   "merge_score", "transform",
                                      "transform_score"};
        py::list out;
        for (const auto& x : t) {
            out.append(
                py::make_tuple(kinds[x.kind >= 0 && x.kind <= 6 ? x.kind : 0], x.level, x.ms));
        }
        return out;
    }
    py::array_t<float> scores_host(const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v,
                                   const PyArrayT<size_t>& widths, bool find_max) {
        const size_t nc = m_plan.info().fold_size_time / (2 * m_cfg.params().nbins);
        const size_t n  = find_max ? nc : nc * static_cast<size_t>(widths.size());
        py::array_t<float> out(static_cast<py::ssize_t>(n));
        check(loki_ffa_scores_host(m_h.get(), ts_e.data(), ts_v.data(),
                                   static_cast<size_t>(ts_e.size()), widths.data(),
                                   static_cast<size_t>(widths.size()), out.mutable_data(), n,
                                   find_max ? 1 : 0));
        return out;
    }

    /// Score-and-filter (score_cuda.cu:382-443): (indices, snr) sorted by index
    /// of every (trial, width) -- or trial, with find_max -- whose S/N exceeds
    /// the threshold.  The output is sized by a first pass; a second pass runs
    /// only when more candidates pass than that first guess holds.
    py::tuple scores_filter_host(const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v,
                                 const PyArrayT<size_t>& widths, double threshold,
                                 bool find_max) {
        const size_t nc  = m_plan.info().fold_size_time / (2 * m_cfg.params().nbins);
        const size_t all = find_max ? nc : nc * static_cast<size_t>(widths.size());
        size_t cap = std::min<size_t>(all, size_t{1} << 22);
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<uint32_t> idx(std::max<size_t>(cap, 1));
            std::vector<float> snr(std::max<size_t>(cap, 1));
            size_t n = 0;
            check(loki_ffa_scores_filter_host(m_h.get(), ts_e.data(), ts_v.data(),
                                              static_cast<size_t>(ts_e.size()), widths.data(),
                                              static_cast<size_t>(widths.size()),
                                              static_cast<float>(threshold), find_max ? 1 : 0,
                                              idx.data(), snr.data(), cap, &n));
            if (n <= cap) {
                idx.resize(n);
                snr.resize(n);
                return py::make_tuple(to_array(idx), to_array(snr));
            }
            cap = n;
        }
        throw std::runtime_error("score filter: candidate count changed between passes");
    }


    int fold_type() const { return m_fold_type; }

private:
    Config m_cfg;
    int m_device;
    int m_fold_type;
    Plan m_plan;
    std::shared_ptr<loki_ffa_s> m_h;
};

/// Distinct C++ type for the complex-fold class (pybind11 binds one Python
/// class per C++ type).
struct FFAComplex : FFA {
    FFAComplex(const Config& cfg, int device_id, const Workspace* ws = nullptr)
        : FFA(cfg, device_id, LOKI_FOLD_FOURIER, ws) {}
};

/// libculoki.ffa.FFABatchCUDA: many frequency-only searches advanced together.
class Batch {
public:
    Batch(const std::vector<Config>& cfgs, std::optional<PyArrayT<size_t>> widths, int device_id)
        : m_cfgs(cfgs) {
        std::vector<loki_config_t> hs;
        for (const auto& c : cfgs) {
            hs.push_back(c.handle());
        }
        std::vector<size_t> w;
        if (widths.has_value()) {
            w.assign(widths->data(), widths->data() + widths->size());
        }
        loki_batch_t h = nullptr;
        check(loki_batch_create(hs.data(), hs.size(), w.empty() ? nullptr : w.data(), w.size(),
                                device_id, &h));
        m_h.reset(h, loki_batch_destroy);
    }
    size_t size() const { return m_cfgs.size(); }
    PlanTime plan(size_t i) const { return PlanTime(m_cfgs.at(i)); }
    void set_input(size_t i, const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v) {
        if (ts_v.size() != ts_e.size()) {
            throw std::runtime_error("FFA::execute: ts_v must have size nsamps");
        }
        check(loki_batch_set_input_host(m_h.get(), i, ts_e.data(), ts_v.data(),
                                        static_cast<size_t>(ts_e.size())));
    }
    void set_input_device(size_t i, uintptr_t ts_e, uintptr_t ts_v, size_t nsamps,
                          uintptr_t stream) {
        check(loki_batch_set_input_device(m_h.get(), i, reinterpret_cast<const float*>(ts_e),
                                          reinterpret_cast<const float*>(ts_v), nsamps,
                                          reinterpret_cast<void*>(stream)));
    }
    py::tuple input_ptrs(size_t i) const {
        float *e = nullptr, *v = nullptr;
        check(loki_batch_series(m_h.get(), i, &e, &v, nullptr, nullptr, nullptr));
        return py::make_tuple(reinterpret_cast<uintptr_t>(e), reinterpret_cast<uintptr_t>(v));
    }
    py::tuple scores_ptr(size_t i) const {
        float* sp = nullptr;
        size_t n = 0;
        check(loki_batch_series(m_h.get(), i, nullptr, nullptr, &sp, &n, nullptr));
        return py::make_tuple(reinterpret_cast<uintptr_t>(sp), n);
    }
    py::array_t<float> scores(size_t i) const {
        size_t n = 0;
        check(loki_batch_series(m_h.get(), i, nullptr, nullptr, nullptr, &n, nullptr));
        py::array_t<float> out(static_cast<py::ssize_t>(n));
        check(loki_batch_scores_host(m_h.get(), i, out.mutable_data(), n));
        return out;
    }
    py::array_t<size_t> widths(size_t i) const {
        size_t n = 0;
        check(loki_batch_widths(m_h.get(), i, nullptr, 0, &n));
        py::array_t<size_t> out(static_cast<py::ssize_t>(n));
        check(loki_batch_widths(m_h.get(), i, out.mutable_data(), n, &n));
        return out;
    }
    void execute(uintptr_t stream) {
        check(loki_batch_execute(m_h.get(), reinterpret_cast<void*>(stream)));
    }
    void set_graph(bool on) { check(loki_batch_set_graph(m_h.get(), on ? 1 : 0)); }
    /// Keys: merge_group, merge_run, final_group, final_run, persistent and
    /// score_lanes_<nbins> (nbins 8..128); keys left out keep their defaults.
    void set_schedule(const std::map<std::string, uint32_t>& sc) {
        loki_batch_schedule b{16, 16, 16, 16, 0, {0, 0, 0, 0, 0}};
        for (const auto& [k, v] : sc) {
            if (k == "merge_group") {
                b.merge_group = v;
            } else if (k == "merge_run") {
                b.merge_run = v;
            } else if (k == "final_group") {
                b.final_group = v;
            } else if (k == "final_run") {
                b.final_run = v;
            } else if (k == "persistent") {
                b.persistent = v;
            } else if (k.rfind("score_lanes_", 0) == 0) {
                const int nb = std::stoi(k.substr(12));
                int idx      = -1;
                for (int i = 0; i < 5; ++i) {
                    idx = (nb == (8 << i)) ? i : idx;
                }
                if (idx < 0) {
                    throw std::runtime_error("unknown schedule key " + k);
                }
                b.score_lanes[idx] = v;
            } else {
                throw std::runtime_error("unknown schedule key " + k);
            }
        }
        check(loki_batch_set_schedule(m_h.get(), &b));
    }
    static std::map<std::string, uint32_t> as_map(const loki_batch_schedule& b) {
        std::map<std::string, uint32_t> m{{"merge_group", b.merge_group},
                                          {"merge_run", b.merge_run},
                                          {"final_group", b.final_group},
                                          {"final_run", b.final_run},
                                          {"persistent", b.persistent}};
        for (int i = 0; i < 5; ++i) {
            m["score_lanes_" + std::to_string(8 << i)] = b.score_lanes[i];
        }
        return m;
    }
    std::map<std::string, uint32_t> autotune(int repeats, uintptr_t stream) {
        loki_batch_schedule b{};
        check(loki_batch_autotune(m_h.get(), reinterpret_cast<void*>(stream), repeats, &b));
        return as_map(b);
    }
    uint32_t errors() {
        uint32_t f = 0;
        check(loki_batch_errors(m_h.get(), &f));
        return f;
    }
    py::dict info() const {
        size_t n = 0, trials = 0, bytes = 0, launches = 0;
        check(loki_batch_info(m_h.get(), &n, &trials, &bytes, &launches));
        py::dict d;
        d["nseries"]      = n;
        d["total_trials"] = trials;
        d["device_bytes"] = bytes;
        d["launches"]     = launches;
        return d;
    }

private:
    std::vector<Config> m_cfgs;
    std::shared_ptr<loki_batch_s> m_h;
};

void check_nsamps(const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v, const Config& cfg) {
    if (static_cast<size_t>(ts_e.size()) != cfg.params().nsamps) {
        throw std::runtime_error("FFA::execute: ts_e must have size nsamps");
    }
    if (ts_v.size() != ts_e.size()) {
        throw std::runtime_error("FFA::execute: ts_v must have size nsamps");
    }
}

}  // namespace

namespace {

// regions.hpp FFARegionStats (bound at loki_python.cpp:482-510).
struct RegionStats {
    loki_region_stats s{};
};

// regions.cpp:163-420 FFARegionPlanner<T> (bound at loki_templates.hpp:76-84).
class RegionPlanner {
public:
    RegionPlanner(const Config& cfg, int fold_type) {
        loki_region_planner_t h = nullptr;
        check(loki_region_planner_create(cfg.handle(), fold_type, 1, &h));
        adopt(h);
    }
    explicit RegionPlanner(loki_region_planner_t h) { adopt(h); }
    std::vector<Config> cfgs() const {
        std::vector<Config> out;
        for (size_t i = 0; i < m_n; ++i) {
            loki_config_t c = nullptr;
            check(loki_region_planner_config(m_h.get(), i, &c));
            out.emplace_back(c);
        }
        return out;
    }
    size_t nregions() const { return m_n; }
    RegionStats stats() const {
        RegionStats st;
        check(loki_region_planner_stats(m_h.get(), &st.s));
        return st;
    }
    py::list chunk_stats() const {
        py::list out;
        for (size_t i = 0; i < m_n; ++i) {
            loki_ffa_chunk_stats c{};
            check(loki_region_planner_chunk(m_h.get(), i, &c));
            py::dict d;
            d["nominal_f_start"]  = c.nominal_f_start;
            d["nominal_f_end"]    = c.nominal_f_end;
            d["actual_f_start"]   = c.actual_f_start;
            d["actual_f_end"]     = c.actual_f_end;
            d["nominal_width"]    = c.nominal_width;
            d["actual_width"]     = c.actual_width;
            d["total_memory_gb"]  = c.total_memory_gb;
            d["overlap_fraction"] = c.overlap_fraction;
            out.append(d);
        }
        return out;
    }

private:
    std::shared_ptr<loki_region_planner_s> m_h;
    size_t m_n = 0;
    void adopt(loki_region_planner_t h) {
        m_h.reset(h, loki_region_planner_destroy);
        check(loki_region_planner_nregions(h, &m_n));
    }
};

// FFAFreqSweepCUDA (ffa_freq_sweep_cuda.cu:21-365, bound at loki_python_cuda.cpp:205-217).
class FreqSweep {
public:
    FreqSweep(const Config& cfg, int device_id) : m_cfg(cfg) {
        loki_freq_sweep_t h = nullptr;
        check(loki_freq_sweep_create(cfg.handle(), device_id, &h));
        m_h.reset(h, loki_freq_sweep_destroy);
    }
    /// Runs every chunk and writes {outdir}/{file_prefix}_ffa_results.npz (the reference writes
    /// the same datasets to {file_prefix}_ffa_results.h5); returns the archive's path.
    py::object execute(const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v,
                       const std::string& outdir, const std::string& file_prefix) {
        if (ts_e.size() != ts_v.size()) {
            throw std::invalid_argument("ts_e and ts_v must have the same size");
        }
        size_t n = 0;
        {
            py::gil_scoped_release nogil;
            check(loki_freq_sweep_execute(m_h.get(), ts_e.data(), ts_v.data(),
                                          static_cast<size_t>(ts_e.size()), &n));
        }
        auto [snr, ps] = results();
        auto names     = m_cfg.param_names();
        names.emplace_back("width");
        auto path = py::module_::import("pathlib").attr("Path")(outdir) /
                    py::str(file_prefix + "_ffa_results.npz");
        py::module_::import("loki_amd.cands")
            .attr("write_results")(path, names, m_cfg.score_widths(), snr, ps);
        return path;
    }
    std::pair<py::array_t<float>, py::array_t<double>> results() const {
        size_t n = 0;
        check(loki_freq_sweep_count(m_h.get(), &n));
        const auto np = static_cast<py::ssize_t>(m_cfg.params().nparams + 1);
        py::array_t<float> snr(static_cast<py::ssize_t>(n));
        py::array_t<double> ps({static_cast<py::ssize_t>(n), np});
        check(loki_freq_sweep_results(m_h.get(), snr.mutable_data(), ps.mutable_data(), n));
        return {snr, ps};
    }
    std::vector<size_t> per_region() const {
        std::vector<size_t> v(planner().nregions());
        check(loki_freq_sweep_per_region(m_h.get(), v.data(), v.size()));
        return v;
    }
    RegionPlanner planner() const {
        loki_region_planner_t p = nullptr;
        check(loki_freq_sweep_planner(m_h.get(), &p));
        return RegionPlanner(p);
    }

private:
    Config m_cfg;
    std::shared_ptr<loki_freq_sweep_s> m_h;
};

}  // namespace

PYBIND11_MODULE(_loki_amd, m) {
    m.doc() = "MI355X (gfx950) FFA hot path: loki-compatible bindings over the loki_amd C-ABI";

    // ---------------------------------------------------------------- configs
    auto m_configs = m.def_submodule("configs", "Configs submodule");
    py::class_<Config>(m_configs, "PulsarSearchConfig")
        .def(py::init<size_t, double, size_t, double, const PyArrayT<double>&, double, double,
                      bool, int, double, double, size_t, size_t, std::optional<size_t>,
                      std::optional<size_t>, double, size_t, size_t, double, double, double,
                      double, bool>(),
             py::arg("nsamps"), py::arg("tsamp"), py::arg("nbins"), py::arg("eta"),
             py::arg("param_limits"), py::arg("ducy_max") = 0.2, py::arg("wtsp") = 1.5,
             py::arg("use_fourier") = true, py::arg("nthreads") = 1,
             py::arg("max_process_memory_gb") = 8.0, py::arg("octave_scale") = 2.0,
             py::arg("nbins_max") = 1024, py::arg("nbins_min_lossy_bf") = 64,
             py::arg("bseg_brute") = std::nullopt, py::arg("bseg_ffa") = std::nullopt,
             py::arg("snr_min") = 5.0, py::arg("max_passing_candidates") = 1U << 22U,
             py::arg("prune_poly_order") = 3, py::arg("p_orb_min") = 1e-5,
             py::arg("m_c_max") = 10.0, py::arg("m_p_min") = 1.4,
             py::arg("minimum_snap_cells") = 5.0, py::arg("use_conservative_tile") = false)
        .def_property_readonly("nsamps", [](const Config& c) { return c.params().nsamps; })
        .def_property_readonly("tsamp", [](const Config& c) { return c.params().tsamp; })
        .def_property_readonly("tobs", [](const Config& c) { return c.info().tobs; })
        .def_property_readonly("nbins", [](const Config& c) { return c.params().nbins; })
        .def_property_readonly("nbins_f", [](const Config& c) { return c.info().nbins_f; })
        .def_property_readonly("eta", [](const Config& c) { return c.params().eta; })
        .def_property_readonly("param_limits", &Config::param_limits)
        .def_property_readonly("ducy_max", [](const Config& c) { return c.params().ducy_max; })
        // not exposed by the reference's bindings: the threshold of the score-and-filter step
        .def_property_readonly("snr_min", [](const Config& c) { return c.params().snr_min; })
        .def_property_readonly("wtsp", [](const Config& c) { return c.params().wtsp; })
        .def_property_readonly("prune_poly_order",
                               [](const Config& c) { return c.params().prune_poly_order; })
        .def_property_readonly("bseg_brute", [](const Config& c) { return c.info().bseg_brute; })
        .def_property_readonly("bseg_ffa", [](const Config& c) { return c.info().bseg_ffa; })
        .def_property_readonly("p_orb_min", [](const Config& c) { return c.params().p_orb_min; })
        .def_property_readonly("minimum_snap_cells",
                               [](const Config& c) { return c.params().minimum_snap_cells; })
        .def_property_readonly("use_fourier",
                               [](const Config& c) { return c.params().use_fourier != 0; })
        .def_property_readonly("use_conservative_tile",
                               [](const Config& c) { return c.params().use_conservative_tile != 0; })
        .def_property_readonly("nthreads", [](const Config& c) { return c.params().nthreads; })
        .def_property_readonly("tseg_brute", [](const Config& c) { return c.info().tseg_brute; })
        .def_property_readonly("tseg_ffa", [](const Config& c) { return c.info().tseg_ffa; })
        .def_property_readonly("niters_ffa", [](const Config& c) { return c.info().niters_ffa; })
        .def_property_readonly("nparams", [](const Config& c) { return c.info().nparams; })
        .def_property_readonly("param_names", &Config::param_names)
        .def_property_readonly("f_min", [](const Config& c) { return c.info().f_min; })
        .def_property_readonly("f_max", [](const Config& c) { return c.info().f_max; })
        .def_property_readonly("score_widths", &Config::score_widths)
        .def_property_readonly("n_scoring_widths",
                               [](const Config& c) { return c.info().n_scoring_widths; })
        .def("dparams_f", [](const Config& c, double t) { return c.dparams(t, 1); },
             py::arg("tseg_cur"))
        .def("dparams", [](const Config& c, double t) { return c.dparams(t, 0); },
             py::arg("tseg_cur"))
        .def("dparams_actual", [](const Config& c, double t) { return c.dparams(t, 2); },
             py::arg("tseg_cur"));

    // ------------------------------------------------------------------ plans
    auto m_plans = m.def_submodule("plans", "Plans submodule");
    py::class_<Plan>(m_plans, "FFAPlanBase")
        .def(py::init<const Config&, int>(), py::arg("cfg"), py::arg("fold_type") = 0)
        .def_property_readonly("n_params", [](const Plan& p) { return p.info().n_params; })
        .def_property_readonly("n_levels", [](const Plan& p) { return p.info().n_levels; })
        .def_property_readonly("segment_lens", [](const Plan& p) {
            return p.level_array([](const loki_plan_level& l) { return l.segment_len; });
        })
        .def_property_readonly("nsegments", [](const Plan& p) {
            return p.level_array([](const loki_plan_level& l) { return l.nsegments; });
        })
        .def_property_readonly("tsegments", [](const Plan& p) {
            return p.level_array([](const loki_plan_level& l) { return l.tsegment; });
        })
        .def_property_readonly("ncoords", [](const Plan& p) {
            return p.level_array([](const loki_plan_level& l) { return l.ncoords; });
        })
        .def_property_readonly("ncoords_lb", [](const Plan& p) {
            return p.level_array([](const loki_plan_level& l) { return l.ncoords_lb; });
        })
        .def_property_readonly("ncoords_offsets", [](const Plan& p) {
            // u32 offsets with the final sentinel (plans.cpp:311-319).
            auto off = p.level_array([](const loki_plan_level& l) {
                             return static_cast<uint32_t>(l.ncoords_offset);
                         }).cast<std::vector<uint32_t>>();
            off.push_back(static_cast<uint32_t>(p.info().coord_size));
            return to_array(off);
        })
        .def_property_readonly("param_counts",
                               [](const Plan& p) { return p.per_level(true, false, false, false); })
        .def_property_readonly("param_cart_strides",
                               [](const Plan& p) { return p.per_level(false, true, false, false); })
        .def_property_readonly("dparams",
                               [](const Plan& p) { return p.per_level(false, false, true, false); })
        .def_property_readonly("dparams_actual",
                               [](const Plan& p) { return p.per_level(false, false, false, true); })
        .def_property_readonly("config", &Plan::config)
        .def_property_readonly("coord_size", [](const Plan& p) { return p.info().coord_size; })
        .def_property_readonly("coord_memory_usage",
                               [](const Plan& p) { return p.info().coord_memory_usage; })
        .def("resolve_coordinates",
             [](const Plan& p, int device_id) { return p.resolve_coordinates(false, device_id); },
             py::arg("device_id") = 0)
        .def("resolve_coordinates_freq",
             [](const Plan& p, int device_id) { return p.resolve_coordinates(true, device_id); },
             py::arg("device_id") = 0)
        .def_property_readonly("param_grid",
                               [](const Plan& p) {
                                   py::list out;
                                   for (size_t l = 0; l < p.info().n_levels; ++l) {
                                       out.append(p.compute_param_grid(l));
                                   }
                                   return out;
                               })
        .def_property_readonly("params_dict", &Plan::params_dict)
        .def("compute_param_grid", &Plan::compute_param_grid, py::arg("ffa_level"))
        .def_property_readonly("fold_shapes", [](const Plan& p) { return p.fold_shapes(false); })
        .def_property_readonly("fold_shapes_time",
                               [](const Plan& p) { return p.fold_shapes(true); })
        .def_property_readonly("brute_fold_size",
                               [](const Plan& p) { return p.info().brute_fold_size; })
        .def_property_readonly("fold_size", [](const Plan& p) { return p.info().fold_size; })
        .def_property_readonly("fold_size_time",
                               [](const Plan& p) { return p.info().fold_size_time; })
        .def_property_readonly("buffer_size", [](const Plan& p) { return p.info().buffer_size; })
        .def_property_readonly("buffer_size_time",
                               [](const Plan& p) { return p.info().buffer_size_time; })
        .def_property_readonly("buffer_memory_usage",
                               [](const Plan& p) { return p.info().buffer_memory_usage; })
        .def("get_gflops", [](const Plan& p, bool t) {
            return t ? p.info().gflops_return_in_time : p.info().gflops;
        }, py::arg("return_in_time") = false);
    py::class_<PlanTime, Plan>(m_plans, "FFAPlanTime").def(py::init<const Config&>(), py::arg("cfg"));
    py::class_<RegionStats>(m_plans, "FFARegionStats")
        .def(py::init([](size_t max_buffer_size, size_t max_coord_size, size_t max_ncoords,
                         size_t max_ffa_levels, size_t n_widths, size_t n_params,
                         size_t n_samps, size_t max_passing_candidates, bool use_fourier,
                         bool use_gpu) {
                 RegionStats st;
                 check(loki_region_stats_compute(max_buffer_size, max_coord_size, max_ncoords,
                                                 max_ffa_levels, n_widths, n_params, n_samps,
                                                 max_passing_candidates, use_fourier ? 1 : 0,
                                                 use_gpu ? 1 : 0, &st.s));
                 return st;
             }),
             py::arg("max_buffer_size"), py::arg("max_coord_size"), py::arg("max_ncoords"),
             py::arg("max_ffa_levels"), py::arg("n_widths"), py::arg("n_params"),
             py::arg("n_samps"), py::arg("max_passing_candidates"), py::arg("use_fourier"),
             py::arg("use_gpu") = false)
        .def_property_readonly("max_buffer_size",
                               [](const RegionStats& r) { return r.s.max_buffer_size; })
        .def_property_readonly("max_coord_size",
                               [](const RegionStats& r) { return r.s.max_coord_size; })
        .def_property_readonly("max_ncoords", [](const RegionStats& r) { return r.s.max_ncoords; })
        .def_property_readonly("max_ffa_levels",
                               [](const RegionStats& r) { return r.s.max_ffa_levels; })
        .def_property_readonly("max_buffer_size_time",
                               [](const RegionStats& r) { return r.s.max_buffer_size_time; })
        .def_property_readonly("max_scores_size",
                               [](const RegionStats& r) { return r.s.max_scores_size; })
        .def_property_readonly("write_param_sets_size",
                               [](const RegionStats& r) { return r.s.write_param_sets_size; })
        .def_property_readonly("buffer_memory_usage",
                               [](const RegionStats& r) { return r.s.buffer_memory_usage; })
        .def_property_readonly("coord_memory_usage",
                               [](const RegionStats& r) { return r.s.coord_memory_usage; })
        .def_property_readonly("extra_memory_usage",
                               [](const RegionStats& r) { return r.s.extra_memory_usage; })
        .def_property_readonly("freq_sweep_memory_usage",
                               [](const RegionStats& r) { return r.s.freq_sweep_memory_usage; });
    py::class_<RegionPlanner>(m_plans, "FFARegionPlanner")
        .def(py::init<const Config&, int>(), py::arg("cfg"), py::arg("fold_type") = 0)
        .def_property_readonly("cfgs", &RegionPlanner::cfgs)
        .def_property_readonly("nregions", &RegionPlanner::nregions)
        .def_property_readonly("stats", &RegionPlanner::stats)
        .def_property_readonly("chunk_stats", &RegionPlanner::chunk_stats);
    m_plans.def(
        "FFARegionPlannerTime",
        [](const Config& cfg) { return RegionPlanner(cfg, LOKI_FOLD_TIME); }, py::arg("cfg"));
    m_plans.def(
        "FFARegionPlannerFourier",
        [](const Config& cfg) { return RegionPlanner(cfg, LOKI_FOLD_FOURIER); }, py::arg("cfg"));
    m_plans.def(
        "generate_ffa_regions",
        [](double p_min, double p_max, double tsamp, size_t nbins_min, double eta_min,
           double growth_factor, std::optional<size_t> nbins_max) {
            size_t n = 0;
            const size_t cap_bins = nbins_max.value_or(0);
            check(loki_generate_ffa_regions(p_min, p_max, tsamp, nbins_min, eta_min,
                                            growth_factor, cap_bins, nullptr, 0, &n));
            std::vector<loki_ffa_region> r(n);
            check(loki_generate_ffa_regions(p_min, p_max, tsamp, nbins_min, eta_min,
                                            growth_factor, cap_bins, r.data(), n, &n));
            // a structured array of coord.hpp FFARegion (registered at loki_python.cpp:380)
            py::list fields;
            fields.append(py::make_tuple("f_start", "<f8"));
            fields.append(py::make_tuple("f_end", "<f8"));
            fields.append(py::make_tuple("nbins", "<u8"));
            fields.append(py::make_tuple("eta", "<f8"));
            py::dtype dt = py::module_::import("numpy").attr("dtype")(fields).cast<py::dtype>();
            py::array out(dt, std::vector<py::ssize_t>{static_cast<py::ssize_t>(n)});
            static_assert(sizeof(loki_ffa_region) == 32, "packed region record");
            std::memcpy(out.mutable_data(), r.data(), n * sizeof(loki_ffa_region));
            return out;
        },
        py::arg("p_min"), py::arg("p_max"), py::arg("tsamp"), py::arg("nbins_min"),
        py::arg("eta_min"), py::arg("growth_factor") = 2.0, py::arg("nbins_max") = py::none());
    py::class_<PlanFourier, Plan>(m_plans, "FFAPlanFourier")
        .def(py::init<const Config&>(), py::arg("cfg"));

    // ----------------------------------------------------------------- scores
    auto m_scores = m.def_submodule("scores", "Scores submodule");
    m_scores.def(
        "generate_box_width_trials",
        [](size_t nbins, double ducy_max, double wtsp) {
            size_t n = 0;
            check(loki_generate_box_width_trials(nbins, ducy_max, wtsp, nullptr, 0, &n));
            std::vector<size_t> v(n);
            check(loki_generate_box_width_trials(nbins, ducy_max, wtsp, v.data(), n, &n));
            return to_array(v);
        },
        py::arg("nbins"), py::arg("ducy_max") = 0.3, py::arg("wtsp") = 1.5);
    m_scores.def(
        "snr_boxcar_3d",
        [](const PyArrayT<float>& arr, const PyArrayT<size_t>& widths, int device_id) {
            if (arr.ndim() != 3 || widths.ndim() != 1) {
                throw std::runtime_error(
                    "Input array must be 3-dimensional, widths must be 1-dimensional");
            }
            if (arr.shape(0) == 0 || arr.shape(1) == 0 || widths.size() == 0) {
                throw std::runtime_error("Input arrays cannot be empty");
            }
            const auto nprof = static_cast<size_t>(arr.shape(0));
            const auto nbins = static_cast<size_t>(arr.shape(2));
            py::array_t<float> out({static_cast<py::ssize_t>(nprof), widths.size()});
            check(loki_snr_boxcar_3d_host(arr.data(), nprof, nbins, widths.data(),
                                          static_cast<size_t>(widths.size()), out.mutable_data(),
                                          0, device_id));
            return out;
        },
        py::arg("arr"), py::arg("widths"), py::arg("device_id") = 0);
    m_scores.def(
        "snr_boxcar_3d_filter",
        [](const PyArrayT<float>& arr, const PyArrayT<size_t>& widths, double threshold,
           bool find_max, int device_id) {
            if (arr.ndim() != 3 || widths.ndim() != 1) {
                throw std::runtime_error(
                    "Input array must be 3-dimensional, widths must be 1-dimensional");
            }
            if (arr.shape(0) == 0 || arr.shape(1) == 0 || widths.size() == 0) {
                throw std::runtime_error("Input arrays cannot be empty");
            }
            const auto nprof = static_cast<size_t>(arr.shape(0));
            const auto nbins = static_cast<size_t>(arr.shape(2));
            const size_t cap = find_max ? nprof : nprof * static_cast<size_t>(widths.size());
            std::vector<uint32_t> idx(std::max<size_t>(cap, 1));
            std::vector<float> snr(std::max<size_t>(cap, 1));
            size_t n = 0;
            check(loki_snr_boxcar_3d_filter_host(arr.data(), nprof, nbins, widths.data(),
                                                 static_cast<size_t>(widths.size()),
                                                 static_cast<float>(threshold), find_max ? 1 : 0,
                                                 idx.data(), snr.data(), cap, &n, device_id));
            idx.resize(n);
            snr.resize(n);
            return py::make_tuple(to_array(idx), to_array(snr));
        },
        py::arg("arr"), py::arg("widths"), py::arg("threshold"), py::arg("find_max") = false,
        py::arg("device_id") = 0);
    m_scores.def(
        "snr_boxcar_3d_max",
        [](const PyArrayT<float>& arr, const PyArrayT<size_t>& widths, int device_id) {
            if (arr.ndim() != 3 || widths.ndim() != 1) {
                throw std::runtime_error(
                    "Input array must be 3-dimensional, widths must be 1-dimensional");
            }
            if (arr.shape(0) == 0 || arr.shape(1) == 0 || widths.size() == 0) {
                throw std::runtime_error("Input arrays cannot be empty");
            }
            const auto nprof = static_cast<size_t>(arr.shape(0));
            const auto nbins = static_cast<size_t>(arr.shape(2));
            py::array_t<float> out(static_cast<py::ssize_t>(nprof));
            check(loki_snr_boxcar_3d_host(arr.data(), nprof, nbins, widths.data(),
                                          static_cast<size_t>(widths.size()), out.mutable_data(),
                                          1, device_id));
            return out;
        },
        py::arg("arr"), py::arg("widths"), py::arg("device_id") = 0);
    m_scores.def(
        "snr_boxcar_2d_max",
        [](const PyArrayT<float>& arr, const PyArrayT<size_t>& widths, float stdnoise,
           int device_id) {
            if (arr.ndim() != 2 || widths.ndim() != 1) {
                throw std::runtime_error(
                    "Input array must be 2-dimensional, widths must be 1-dimensional");
            }
            if (arr.shape(0) == 0 || arr.shape(1) == 0 || widths.size() == 0) {
                throw std::runtime_error("Input arrays cannot be empty");
            }
            const auto nprof = static_cast<size_t>(arr.shape(0));
            const auto nbins = static_cast<size_t>(arr.shape(1));
            py::array_t<float> out(static_cast<py::ssize_t>(nprof));
            check(loki_snr_boxcar_2d_max_host(arr.data(), nprof, nbins, widths.data(),
                                              static_cast<size_t>(widths.size()), stdnoise,
                                              out.mutable_data(), device_id));
            return out;
        },
        py::arg("arr"), py::arg("widths"), py::arg("stdnoise") = 1.0F, py::arg("device_id") = 0);

    // ------------------------------------------------------------------- fold
    auto m_fold = m.def_submodule("fold", "Fold submodule");
    m_fold.def(
        "compute_brute_fold_time_cuda",
        [](const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v,
           const PyArrayT<double>& freq_arr, size_t segment_len, size_t nbins, double tsamp,
           double t_ref, int device_id) {
            if (ts_v.size() != ts_e.size()) {
                throw std::runtime_error("BruteFold::Impl::execute: ts_v must have size nsamps");
            }
            const auto nsamps = static_cast<size_t>(ts_e.size());
            if (segment_len == 0 || nsamps % segment_len != 0) {
                throw std::runtime_error(
                    "BruteFold::Impl: Number of samples is not a multiple of segment length");
            }
            const size_t n = (nsamps / segment_len) * static_cast<size_t>(freq_arr.size()) * 2 * nbins;
            py::array_t<float> out(static_cast<py::ssize_t>(n));
            check(loki_brute_fold_time_host(ts_e.data(), ts_v.data(), nsamps, freq_arr.data(),
                                            static_cast<size_t>(freq_arr.size()), segment_len,
                                            nbins, tsamp, t_ref, out.mutable_data(), device_id));
            return out;
        },
        py::arg("ts_e"), py::arg("ts_v"), py::arg("freq_arr"), py::arg("segment_len"),
        py::arg("nbins"), py::arg("tsamp"), py::arg("t_ref") = 0.0F, py::arg("device_id") = 0);

    m_fold.def(
        "compute_brute_fold_fourier_cuda",
        [](const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v,
           const PyArrayT<double>& freq_arr, size_t segment_len, size_t nbins, double tsamp,
           double t_ref, int device_id) {
            if (ts_v.size() != ts_e.size()) {
                throw std::runtime_error("BruteFold::Impl::execute: ts_v must have size nsamps");
            }
            const auto nsamps = static_cast<size_t>(ts_e.size());
            if (segment_len == 0 || nsamps % segment_len != 0) {
                throw std::runtime_error(
                    "BruteFold::Impl: Number of samples is not a multiple of segment length");
            }
            const size_t n = (nsamps / segment_len) * static_cast<size_t>(freq_arr.size()) * 2 *
                             (nbins / 2 + 1);
            py::array_t<std::complex<float>> out(static_cast<py::ssize_t>(n));
            check(loki_brute_fold_fourier_host(
                ts_e.data(), ts_v.data(), nsamps, freq_arr.data(),
                static_cast<size_t>(freq_arr.size()), segment_len, nbins, tsamp, t_ref,
                reinterpret_cast<float*>(out.mutable_data()), device_id));
            return out;
        },
        py::arg("ts_e"), py::arg("ts_v"), py::arg("freq_arr"), py::arg("segment_len"),
        py::arg("nbins"), py::arg("tsamp"), py::arg("t_ref") = 0.0F, py::arg("device_id") = 0);

    // -------------------------------------------------------------------- ffa
    auto m_ffa = m.def_submodule("ffa", "FFA submodule");
    py::class_<Workspace>(m_ffa, "FFAWorkspaceCUDA")
        .def(py::init<int, size_t, size_t>(), py::arg("device_id") = 0,
             py::arg("buffer_bytes") = 0, py::arg("coord_bytes") = 0);
    py::class_<FFA>(m_ffa, "FFACUDA")
        .def(py::init([](const Config& cfg, int device_id, const Workspace* ws) {
                 return std::make_unique<FFA>(cfg, device_id, LOKI_FOLD_TIME, ws);
             }),
             py::arg("cfg"), py::arg("device_id") = 0, py::arg("workspace") = nullptr,
             py::keep_alive<1, 4>())
        .def_property_readonly("plan", &FFA::plan)
        .def("execute", &FFA::execute, py::arg("ts_e"), py::arg("ts_v"), py::arg("fold"))
        .def("execute_device", &FFA::execute_device, py::arg("ts_e_ptr"), py::arg("ts_v_ptr"),
             py::arg("fold_ptr"), py::arg("fold_len"), py::arg("stream") = 0)
        .def("scores_device", &FFA::scores_device, py::arg("ts_e_ptr"), py::arg("ts_v_ptr"),
             py::arg("widths"), py::arg("scores_ptr"), py::arg("scores_len"),
             py::arg("find_max") = false, py::arg("stream") = 0)
        .def("scores", &FFA::scores_host, py::arg("ts_e"), py::arg("ts_v"), py::arg("widths"),
             py::arg("find_max") = false)
        .def("set_timing", &FFA::set_timing, py::arg("enable"))
        .def("set_fused", &FFA::set_fused, py::arg("enable"))
        .def("set_fuse_depth", &FFA::set_fuse_depth, py::arg("depth"))
        .def("set_two_level_merges", &FFA::set_two_level_merges, py::arg("enable"))
        .def("set_exact_norm", &FFA::set_exact_norm, py::arg("exact"))
        .def("set_fold_storage", &FFA::set_fold_storage, py::arg("storage"))
        .def("scores_filter", &FFA::scores_filter_host, py::arg("ts_e"), py::arg("ts_v"),
             py::arg("widths"), py::arg("threshold"), py::arg("find_max") = false)
        .def("scores_filter_device", &FFA::scores_filter_device, py::arg("ts_e_ptr"),
             py::arg("ts_v_ptr"), py::arg("widths"), py::arg("threshold"), py::arg("find_max"),
             py::arg("indices_ptr"), py::arg("scores_ptr"), py::arg("capacity"),
             py::arg("count_ptr"), py::arg("stream") = 0)
        .def("timings", &FFA::timings);
    py::class_<FFAComplex>(m_ffa, "FFACOMPLEXCUDA")
        .def(py::init<const Config&, int, const Workspace*>(), py::arg("cfg"),
             py::arg("device_id") = 0, py::arg("workspace") = nullptr, py::keep_alive<1, 4>())
        .def_property_readonly("plan", &FFA::plan)
        .def("execute", &FFA::execute_complex, py::arg("ts_e"), py::arg("ts_v"), py::arg("fold"))
        .def("execute", &FFA::execute, py::arg("ts_e"), py::arg("ts_v"), py::arg("fold"))
        .def("execute_device", &FFA::execute_device, py::arg("ts_e_ptr"), py::arg("ts_v_ptr"),
             py::arg("fold_ptr"), py::arg("fold_len"), py::arg("stream") = 0)
        .def("execute_time_device", &FFA::execute_time_device, py::arg("ts_e_ptr"),
             py::arg("ts_v_ptr"), py::arg("fold_ptr"), py::arg("fold_len"),
             py::arg("stream") = 0)
        .def("scores_device", &FFA::scores_device, py::arg("ts_e_ptr"), py::arg("ts_v_ptr"),
             py::arg("widths"), py::arg("scores_ptr"), py::arg("scores_len"),
             py::arg("find_max") = false, py::arg("stream") = 0)
        .def("scores", &FFA::scores_host, py::arg("ts_e"), py::arg("ts_v"), py::arg("widths"),
             py::arg("find_max") = false)
        .def("set_timing", &FFA::set_timing, py::arg("enable"))
        .def("set_fused", &FFA::set_fused, py::arg("enable"))
        .def("set_exact_norm", &FFA::set_exact_norm, py::arg("exact"))
        .def("set_dft_precision", &FFA::set_dft_precision, py::arg("precision"))
        .def("scores_filter", &FFA::scores_filter_host, py::arg("ts_e"), py::arg("ts_v"),
             py::arg("widths"), py::arg("threshold"), py::arg("find_max") = false)
        .def("scores_filter_device", &FFA::scores_filter_device, py::arg("ts_e_ptr"),
             py::arg("ts_v_ptr"), py::arg("widths"), py::arg("threshold"), py::arg("find_max"),
             py::arg("indices_ptr"), py::arg("scores_ptr"), py::arg("capacity"),
             py::arg("count_ptr"), py::arg("stream") = 0)
        .def("timings", &FFA::timings);

    m_ffa.def(
        "compute_ffa_time_cuda",
        [](const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v, const Config& cfg,
           int device_id, bool /*quiet*/, const std::string& fold_storage) {
            check_nsamps(ts_e, ts_v, cfg);
            FFA ffa(cfg, device_id, LOKI_FOLD_TIME);
            ffa.set_fold_storage(fold_storage);
            const size_t buf = ffa.plan().info().buffer_size;
            PyArrayT<float> fold(static_cast<py::ssize_t>(buf));
            ffa.execute(ts_e, ts_v, fold);
            fold.resize({static_cast<py::ssize_t>(ffa.plan().info().fold_size)});
            return py::make_tuple(fold, PlanTime(cfg));
        },
        py::arg("ts_e"), py::arg("ts_v"), py::arg("cfg"), py::arg("device_id") = 0,
        py::arg("quiet") = false, py::arg("fold_storage") = "f32");
    py::class_<FreqSweep>(m_ffa, "FFAFreqSweepCUDA")
        .def(py::init<const Config&, int>(), py::arg("cfg"), py::arg("device_id") = 0)
        .def("execute", &FreqSweep::execute, py::arg("ts_e"), py::arg("ts_v"),
             py::arg("outdir"), py::arg("file_prefix") = "test")
        .def("results", &FreqSweep::results)
        .def_property_readonly("per_region", &FreqSweep::per_region)
        .def_property_readonly("planner", &FreqSweep::planner);
    py::class_<Batch>(m_ffa, "FFABatchCUDA")
        .def(py::init<const std::vector<Config>&, std::optional<PyArrayT<size_t>>, int>(),
             py::arg("cfgs"), py::arg("widths") = std::nullopt, py::arg("device_id") = 0)
        .def_property_readonly("size", &Batch::size)
        .def("plan", &Batch::plan, py::arg("i"))
        .def("set_input", &Batch::set_input, py::arg("i"), py::arg("ts_e"), py::arg("ts_v"))
        .def("set_input_device", &Batch::set_input_device, py::arg("i"), py::arg("ts_e_ptr"),
             py::arg("ts_v_ptr"), py::arg("nsamps"), py::arg("stream") = 0)
        .def("input_ptrs", &Batch::input_ptrs, py::arg("i"))
        .def("scores_ptr", &Batch::scores_ptr, py::arg("i"))
        .def("scores", &Batch::scores, py::arg("i"))
        .def("widths", &Batch::widths, py::arg("i"))
        .def("execute", &Batch::execute, py::arg("stream") = 0)
        .def("set_graph", &Batch::set_graph, py::arg("enable"))
        .def("set_schedule", &Batch::set_schedule, py::arg("schedule"))
        .def("autotune", &Batch::autotune, py::arg("repeats") = 3, py::arg("stream") = 0)
        .def("errors", &Batch::errors)
        .def("info", &Batch::info);
    m_ffa.def(
        "compute_ffa_fourier_cuda",
        [](const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v, const Config& cfg,
           int device_id, bool /*quiet*/) {
            check_nsamps(ts_e, ts_v, cfg);
            FFA ffa(cfg, device_id, LOKI_FOLD_FOURIER);
            const size_t buf = ffa.plan().info().buffer_size;
            PyArrayT<std::complex<float>> fold(static_cast<py::ssize_t>(buf));
            ffa.execute_complex(ts_e, ts_v, fold);
            fold.resize({static_cast<py::ssize_t>(ffa.plan().info().fold_size)});
            return py::make_tuple(fold, PlanFourier(cfg));
        },
        py::arg("ts_e"), py::arg("ts_v"), py::arg("cfg"), py::arg("device_id") = 0,
        py::arg("quiet") = false);
    m_ffa.def(
        "compute_ffa_fourier_return_to_time_cuda",
        [](const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v, const Config& cfg,
           int device_id, bool /*quiet*/) {
            check_nsamps(ts_e, ts_v, cfg);
            FFA ffa(cfg, device_id, LOKI_FOLD_FOURIER);
            const size_t buf = ffa.plan().info().buffer_size_time;
            PyArrayT<float> fold(static_cast<py::ssize_t>(buf));
            ffa.execute(ts_e, ts_v, fold);
            fold.resize({static_cast<py::ssize_t>(ffa.plan().info().fold_size_time)});
            return py::make_tuple(fold, PlanTime(cfg));
        },
        py::arg("ts_e"), py::arg("ts_v"), py::arg("cfg"), py::arg("device_id") = 0,
        py::arg("quiet") = false);
    m_ffa.def(
        "compute_ffa_scores_cuda",
        [](const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v, const Config& cfg,
           int device_id, bool /*quiet*/, std::optional<PyArrayT<size_t>> widths,
           const std::string& fold_storage) {
            check_nsamps(ts_e, ts_v, cfg);
            // ffa_cuda.cu:676-680: Fourier FFA returned to the time domain, or the time FFA
            FFA ffa(cfg, device_id,
                    cfg.params().use_fourier != 0 ? LOKI_FOLD_FOURIER : LOKI_FOLD_TIME);
            ffa.set_fold_storage(fold_storage);
            PyArrayT<size_t> w = widths.has_value() ? *widths : PyArrayT<size_t>(cfg.score_widths());
            auto scores = ffa.scores_host(ts_e, ts_v, w, false);
            return py::make_tuple(scores, PlanTime(cfg));
        },
        py::arg("ts_e"), py::arg("ts_v"), py::arg("cfg"), py::arg("device_id") = 0,
        py::arg("quiet") = false, py::arg("widths") = std::nullopt,
        py::arg("fold_storage") = "f32");

    m_ffa.def(
        "compute_ffa_scores_filter_cuda",
        [](const PyArrayT<float>& ts_e, const PyArrayT<float>& ts_v, const Config& cfg,
           std::optional<double> threshold, std::optional<PyArrayT<size_t>> widths,
           bool find_max, int device_id) {
            // the FFAFreqSweepCUDA region step (ffa_freq_sweep_cuda.cu:219-260): FFA, then
            // score_and_filter_cuda_d at cfg.snr_min
            check_nsamps(ts_e, ts_v, cfg);
            FFA ffa(cfg, device_id,
                    cfg.params().use_fourier != 0 ? LOKI_FOLD_FOURIER : LOKI_FOLD_TIME);
            PyArrayT<size_t> w = widths.has_value() ? *widths : PyArrayT<size_t>(cfg.score_widths());
            auto res = ffa.scores_filter_host(ts_e, ts_v, w,
                                              threshold.value_or(cfg.params().snr_min), find_max);
            return py::make_tuple(res[0], res[1], PlanTime(cfg));
        },
        py::arg("ts_e"), py::arg("ts_v"), py::arg("cfg"), py::arg("threshold") = std::nullopt,
        py::arg("widths") = std::nullopt, py::arg("find_max") = false, py::arg("device_id") = 0);

    m.def("device_count", [] {
        int n = 0;
        check(loki_device_count(&n));
        return n;
    });
}
