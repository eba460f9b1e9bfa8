// extern "C" surface declared in include/loki_amd.h.  Exceptions from the
// C++ layer become status codes plus a thread-local message.
#include "../../include/loki_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "batch.hpp"
#include "ffa.hpp"
#include "kernels.hpp"
#include "regions.hpp"
#include "sweep.hpp"

struct loki_config_s {
    loki_amd::SearchConfig cfg;
};
struct loki_plan_s {
    loki_amd::FFAPlan plan;
};
struct loki_ffa_s {
    loki_amd::FFAExecutor exec;
    loki_plan_s plan_view;
};
struct loki_ffa_workspace_s {
    std::shared_ptr<loki_amd::FFAWorkspace> ws;
    int device;
};
struct loki_region_planner_s {
    loki_amd::FFARegionPlanner planner;
};
struct loki_freq_sweep_s {
    loki_amd::FFAFreqSweep sweep;
};
struct loki_batch_s {
    std::unique_ptr<loki_amd::FFABatch> batch;
    std::vector<loki_plan_s> plan_views;
};

namespace {

thread_local std::string g_last_error;

template <typename F> int guarded(F&& fn) {
    try {
        fn();
        g_last_error.clear();
        return LOKI_OK;
    } catch (const loki_amd::InvalidArgument& e) {
        g_last_error = e.what();
        return LOKI_ERR_INVALID;
    } catch (const loki_amd::HipError& e) {
        g_last_error = e.what();
        return LOKI_ERR_HIP;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return LOKI_ERR_INTERNAL;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return LOKI_ERR_INTERNAL;
    } catch (...) {
        g_last_error = "unknown error";
        return LOKI_ERR_INTERNAL;
    }
}

#define REQ(cond, msg) LOKI_REQUIRE(cond, msg)

loki_amd::FoldKind fold_kind(int t) {
    REQ(t == LOKI_FOLD_TIME || t == LOKI_FOLD_FOURIER, "fold_type must be 0 (time) or 1 (fourier)");
    return t == LOKI_FOLD_TIME ? loki_amd::FoldKind::kTime : loki_amd::FoldKind::kFourier;
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

const char* loki_last_error(void) { return g_last_error.c_str(); }

int loki_device_count(int* count) {
    return guarded([&] {
        REQ(count != nullptr, "count is null");
        loki_amd::hip_check(hipGetDeviceCount(count), "hipGetDeviceCount");
    });
}

int loki_config_create(const loki_config_params* p, loki_config_t* out) {
    return guarded([&] {
        REQ(p != nullptr && out != nullptr, "null argument");
        REQ(p->nparams >= 1 && p->param_limits != nullptr, "param_limits must be non-empty");
        loki_amd::SearchConfigArgs a;
        a.nsamps = p->nsamps;
        a.tsamp  = p->tsamp;
        a.nbins  = p->nbins;
        a.eta    = p->eta;
        for (size_t i = 0; i < p->nparams; ++i) {
            a.param_limits.push_back({p->param_limits[2 * i], p->param_limits[2 * i + 1]});
        }
        a.ducy_max               = p->ducy_max;
        a.wtsp                   = p->wtsp;
        a.use_fourier            = p->use_fourier != 0;
        a.nthreads               = p->nthreads;
        a.max_process_memory_gb  = p->max_process_memory_gb;
        a.octave_scale           = p->octave_scale;
        a.nbins_max              = p->nbins_max;
        a.nbins_min_lossy_bf     = p->nbins_min_lossy_bf;
        if (p->bseg_brute != 0) {
            a.bseg_brute = p->bseg_brute;
        }
        if (p->bseg_ffa != 0) {
            a.bseg_ffa = p->bseg_ffa;
        }
        a.snr_min                = p->snr_min;
        a.max_passing_candidates = p->max_passing_candidates;
        a.prune_poly_order       = p->prune_poly_order;
        a.p_orb_min              = p->p_orb_min;
        a.m_c_max                = p->m_c_max;
        a.m_p_min                = p->m_p_min;
        a.minimum_snap_cells     = p->minimum_snap_cells;
        a.use_conservative_tile  = p->use_conservative_tile != 0;
        *out = new loki_config_s{loki_amd::SearchConfig(std::move(a))};
    });
}

void loki_config_destroy(loki_config_t cfg) { delete cfg; }

int loki_config_info_get(loki_config_t c, loki_config_info* out) {
    return guarded([&] {
        REQ(c != nullptr && out != nullptr, "null argument");
        const auto& g          = c->cfg;
        out->nbins_f           = g.get_nbins_f();
        out->nparams           = g.get_nparams();
        out->bseg_brute        = g.get_bseg_brute();
        out->bseg_ffa          = g.get_bseg_ffa();
        out->niters_ffa        = g.get_niters_ffa();
        out->n_scoring_widths  = g.get_scoring_widths().size();
        out->tobs              = g.get_tobs();
        out->tseg_brute        = g.get_tseg_brute();
        out->tseg_ffa          = g.get_tseg_ffa();
        out->f_min             = g.get_f_min();
        out->f_max             = g.get_f_max();
    });
}

int loki_config_params_get(loki_config_t c, loki_config_params* out) {
    return guarded([&] {
        REQ(c != nullptr && out != nullptr, "null argument");
        const auto& a               = c->cfg.args();
        out->nsamps                 = a.nsamps;
        out->tsamp                  = a.tsamp;
        out->nbins                  = a.nbins;
        out->eta                    = a.eta;
        out->param_limits           = reinterpret_cast<const double*>(a.param_limits.data());
        out->nparams                = a.param_limits.size();
        out->ducy_max               = a.ducy_max;
        out->wtsp                   = a.wtsp;
        out->use_fourier            = a.use_fourier ? 1 : 0;
        out->nthreads               = a.nthreads;
        out->max_process_memory_gb  = a.max_process_memory_gb;
        out->octave_scale           = a.octave_scale;
        out->nbins_max              = a.nbins_max;
        out->nbins_min_lossy_bf     = a.nbins_min_lossy_bf;
        out->bseg_brute             = c->cfg.get_bseg_brute();
        out->bseg_ffa               = c->cfg.get_bseg_ffa();
        out->snr_min                = a.snr_min;
        out->max_passing_candidates = a.max_passing_candidates;
        out->prune_poly_order       = a.prune_poly_order;
        out->p_orb_min              = a.p_orb_min;
        out->m_c_max                = a.m_c_max;
        out->m_p_min                = a.m_p_min;
        out->minimum_snap_cells     = a.minimum_snap_cells;
        out->use_conservative_tile  = a.use_conservative_tile ? 1 : 0;
    });
}

int loki_config_scoring_widths(loki_config_t c, size_t* out, size_t cap, size_t* n) {
    return guarded([&] {
        REQ(c != nullptr && n != nullptr, "null argument");
        const auto& w = c->cfg.get_scoring_widths();
        *n            = w.size();
        REQ(out == nullptr || cap >= w.size(), "output buffer too small");
        if (out != nullptr) {
            std::memcpy(out, w.data(), w.size() * sizeof(size_t));
        }
    });
}

int loki_config_dparams(loki_config_t c, double tseg_cur, int kind, double* out) {
    return guarded([&] {
        REQ(c != nullptr && out != nullptr, "null argument");
        std::vector<double> v;
        if (kind == 0) {
            v = c->cfg.get_dparams(tseg_cur);
        } else if (kind == 1) {
            v = c->cfg.get_dparams_f(tseg_cur);
        } else if (kind == 2) {
            v = c->cfg.get_dparams_actual(tseg_cur);
        } else {
            REQ(false, "kind must be 0, 1 or 2");
        }
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    });
}

int loki_config_param_grid_count(loki_config_t c, double tseg_cur, size_t* out) {
    return guarded([&] {
        REQ(c != nullptr && out != nullptr, "null argument");
        const auto v = c->cfg.get_param_grid_count(tseg_cur);
        std::memcpy(out, v.data(), v.size() * sizeof(size_t));
    });
}

int loki_generate_box_width_trials(size_t nbins, double ducy_max, double wtsp, size_t* out,
                                   size_t cap, size_t* n) {
    return guarded([&] {
        REQ(n != nullptr, "null argument");
        const auto w = loki_amd::generate_box_width_trials(nbins, ducy_max, wtsp);
        *n           = w.size();
        REQ(out == nullptr || cap >= w.size(), "output buffer too small");
        if (out != nullptr) {
            std::memcpy(out, w.data(), w.size() * sizeof(size_t));
        }
    });
}

int loki_plan_create(loki_config_t cfg, int fold_type, loki_plan_t* out) {
    return guarded([&] {
        REQ(cfg != nullptr && out != nullptr, "null argument");
        *out = new loki_plan_s{loki_amd::FFAPlan(cfg->cfg, fold_kind(fold_type))};
    });
}

void loki_plan_destroy(loki_plan_t plan) { delete plan; }

int loki_plan_info_get(loki_plan_t p, loki_plan_info* out) {
    return guarded([&] {
        REQ(p != nullptr && out != nullptr, "null argument");
        const auto& pl              = p->plan;
        out->n_params               = pl.n_params();
        out->n_levels               = pl.n_levels();
        out->brute_fold_size        = pl.brute_fold_size();
        out->fold_size              = pl.fold_size();
        out->fold_size_time         = pl.fold_size_time();
        out->buffer_size            = pl.buffer_size();
        out->buffer_size_time       = pl.buffer_size_time();
        out->coord_size             = pl.coord_size();
        out->buffer_memory_usage    = pl.buffer_memory_usage();
        out->coord_memory_usage     = pl.coord_memory_usage();
        out->gflops                 = pl.gflops(false);
        out->gflops_return_in_time  = pl.gflops(true);
    });
}

int loki_plan_level_get(loki_plan_t p, size_t l, loki_plan_level* out) {
    return guarded([&] {
        REQ(p != nullptr && out != nullptr, "null argument");
        const auto& pl = p->plan;
        REQ(l < pl.n_levels(), "level out of range");
        std::memset(out, 0, sizeof(*out));
        out->segment_len    = pl.segment_lens()[l];
        out->nsegments      = pl.nsegments()[l];
        out->tsegment       = pl.tsegments()[l];
        out->ncoords        = pl.ncoords()[l];
        out->ncoords_offset = pl.ncoords_offsets64()[l];
        out->ncoords_lb     = pl.ncoords_lb()[l];
        for (size_t i = 0; i < pl.n_params(); ++i) {
            out->param_counts[i]       = pl.param_counts()[l][i];
            out->param_cart_strides[i] = pl.param_cart_strides()[l][i];
            out->dparams[i]            = pl.dparams()[l][i];
            out->dparams_actual[i]     = pl.dparams_actual()[l][i];
        }
        out->fold_ndim = pl.fold_shapes()[l].size();
        for (size_t i = 0; i < out->fold_ndim; ++i) {
            out->fold_shape[i]      = pl.fold_shapes()[l][i];
            out->fold_shape_time[i] = pl.fold_shapes_time()[l][i];
        }
    });
}

int loki_plan_param_grid(loki_plan_t p, size_t level, size_t param, double* out, size_t cap,
                         size_t* n) {
    return guarded([&] {
        REQ(p != nullptr && n != nullptr, "null argument");
        const auto grid = p->plan.compute_param_grid(level);
        REQ(param < grid.size(), "param out of range");
        *n = grid[param].size();
        REQ(out == nullptr || cap >= grid[param].size(), "output buffer too small");
        if (out != nullptr) {
            std::memcpy(out, grid[param].data(), grid[param].size() * sizeof(double));
        }
    });
}

int loki_plan_resolve_coordinates_host(loki_plan_t p, void* coords, size_t coords_bytes,
                                       int device_id) {
    return guarded([&] {
        REQ(p != nullptr && coords != nullptr, "null argument");
        const auto& pl       = p->plan;
        const size_t rec     = pl.n_params() == 1 ? sizeof(loki_amd::FFACoordFreq)
                                                  : sizeof(loki_amd::FFACoord);
        const size_t need    = pl.coord_size() * rec;
        REQ(coords_bytes >= need, "coords must have size >= coord_size records");
        loki_amd::DeviceGuard g(device_id);
        loki_amd::FFAExecutor exec(pl.config(), loki_amd::FoldKind::kTime, device_id);
        loki_amd::DeviceBuffer d(need);
        exec.resolve_coordinates(d.get(), nullptr);
        loki_amd::hip_check(hipDeviceSynchronize(), "resolve");
        const size_t l0 = pl.ncoords()[0] * rec;
        std::memset(coords, 0, l0);
        loki_amd::hip_check(hipMemcpy(static_cast<char*>(coords) + l0,
                                      static_cast<char*>(d.get()) + l0, need - l0,
                                      hipMemcpyDeviceToHost),
                            "copy coords");
    });
}

int loki_ffa_create(loki_config_t cfg, int fold_type, int device_id, loki_ffa_t* out) {
    return guarded([&] {
        REQ(cfg != nullptr && out != nullptr, "null argument");
        auto kind = fold_kind(fold_type);
        loki_amd::FFAExecutor exec(cfg->cfg, kind, device_id);
        loki_amd::FFAPlan plan(cfg->cfg, kind);
        *out = new loki_ffa_s{std::move(exec), loki_plan_s{std::move(plan)}};
    });
}

void loki_ffa_destroy(loki_ffa_t ffa) { delete ffa; }

int loki_ffa_workspace_create(int device_id, size_t buffer_bytes, size_t coord_bytes,
                              loki_ffa_workspace_t* out) {
    return guarded([&] {
        REQ(out != nullptr, "null argument");
        loki_amd::DeviceGuard g(device_id);
        auto ws = std::make_shared<loki_amd::FFAWorkspace>();
        if (buffer_bytes > 0) {
            ws->internal.reserve(buffer_bytes);
            ws->result.reserve(buffer_bytes);
        }
        if (coord_bytes > 0) {
            ws->coords.reserve(coord_bytes);
        }
        *out = new loki_ffa_workspace_s{std::move(ws), device_id};
    });
}

void loki_ffa_workspace_destroy(loki_ffa_workspace_t ws) { delete ws; }

int loki_ffa_create_with_workspace(loki_config_t cfg, int fold_type, loki_ffa_workspace_t ws,
                                   int device_id, loki_ffa_t* out) {
    return guarded([&] {
        REQ(cfg != nullptr && ws != nullptr && out != nullptr, "null argument");
        REQ(ws->device == device_id, "workspace belongs to another device");
        auto kind = fold_kind(fold_type);
        loki_amd::FFAExecutor exec(cfg->cfg, kind, device_id, ws->ws);
        loki_amd::FFAPlan plan(cfg->cfg, kind);
        *out = new loki_ffa_s{std::move(exec), loki_plan_s{std::move(plan)}};
    });
}

int loki_ffa_plan(loki_ffa_t ffa, loki_plan_t* out) {
    return guarded([&] {
        REQ(ffa != nullptr && out != nullptr, "null argument");
        *out = &ffa->plan_view;
    });
}

int loki_ffa_execute_device(loki_ffa_t ffa, const float* ts_e, const float* ts_v, void* fold,
                            size_t fold_len, void* stream) {
    return guarded([&] {
        REQ(ffa != nullptr, "null ffa");
        ffa->exec.execute(ts_e, ts_v, fold, fold_len, as_stream(stream));
    });
}

int loki_ffa_execute_host(loki_ffa_t ffa, const float* ts_e, const float* ts_v, size_t nsamps,
                          void* fold, size_t fold_len) {
    return guarded([&] {
        REQ(ffa != nullptr && ts_e != nullptr && ts_v != nullptr && fold != nullptr,
            "null argument");
        const auto& pl = ffa->exec.plan();
        REQ(nsamps == pl.config().get_nsamps(), "FFA::execute: ts_e must have size nsamps");
        REQ(fold_len == pl.buffer_size(), "FFA::execute: fold must have size buffer_size");
        const size_t elem = pl.kind() == loki_amd::FoldKind::kFourier ? 8 : 4;
        loki_amd::DeviceGuard g(ffa->exec.device_id());
        loki_amd::DeviceBuffer de(nsamps * sizeof(float)), dv(nsamps * sizeof(float));
        loki_amd::DeviceBuffer df(fold_len * elem);
        loki_amd::hip_check(hipMemcpy(de.get(), ts_e, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::hip_check(hipMemcpy(dv.get(), ts_v, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        ffa->exec.execute(de.as<float>(), dv.as<float>(), df.get(), fold_len, nullptr);
        loki_amd::hip_check(hipMemcpy(fold, df.get(), pl.fold_size() * elem, hipMemcpyDeviceToHost),
                            "D2H");
    });
}

int loki_ffa_execute_time_device(loki_ffa_t ffa, const float* ts_e, const float* ts_v,
                                 float* fold, size_t fold_len, void* stream) {
    return guarded([&] {
        REQ(ffa != nullptr, "null ffa");
        ffa->exec.execute_time(ts_e, ts_v, fold, fold_len, as_stream(stream));
    });
}

int loki_ffa_execute_time_host(loki_ffa_t ffa, const float* ts_e, const float* ts_v,
                               size_t nsamps, float* fold, size_t fold_len) {
    return guarded([&] {
        REQ(ffa != nullptr && ts_e != nullptr && ts_v != nullptr && fold != nullptr,
            "null argument");
        const auto& pl = ffa->exec.plan();
        REQ(nsamps == pl.config().get_nsamps(), "FFA::execute: ts_e must have size nsamps");
        REQ(fold_len == 2 * pl.buffer_size(), "FFA::execute: fold must have size 2*buffer_size");
        loki_amd::DeviceGuard g(ffa->exec.device_id());
        loki_amd::DeviceBuffer de(nsamps * 4), dv(nsamps * 4), df(fold_len * 4);
        loki_amd::hip_check(hipMemcpy(de.get(), ts_e, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::hip_check(hipMemcpy(dv.get(), ts_v, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        ffa->exec.execute_time(de.as<float>(), dv.as<float>(), df.as<float>(), fold_len, nullptr);
        loki_amd::hip_check(
            hipMemcpy(fold, df.get(), pl.fold_size_time() * 4, hipMemcpyDeviceToHost), "D2H");
    });
}

int loki_ffa_scores_device(loki_ffa_t ffa, const float* ts_e, const float* ts_v,
                           const size_t* widths, size_t nwidths, float* scores,
                           size_t scores_len, int find_max, void* stream) {
    return guarded([&] {
        REQ(ffa != nullptr && widths != nullptr, "null argument");
        ffa->exec.execute_scores(ts_e, ts_v, std::span<const size_t>(widths, nwidths), scores,
                                 scores_len, find_max != 0, as_stream(stream));
    });
}

int loki_ffa_scores_host(loki_ffa_t ffa, const float* ts_e, const float* ts_v, size_t nsamps,
                         const size_t* widths, size_t nwidths, float* scores, size_t scores_len,
                         int find_max) {
    return guarded([&] {
        REQ(ffa != nullptr && ts_e != nullptr && ts_v != nullptr && scores != nullptr &&
                widths != nullptr,
            "null argument");
        const auto& pl = ffa->exec.plan();
        REQ(nsamps == pl.config().get_nsamps(), "FFA::execute: ts_e must have size nsamps");
        loki_amd::DeviceGuard g(ffa->exec.device_id());
        loki_amd::DeviceBuffer de(nsamps * 4), dv(nsamps * 4), ds(scores_len * 4);
        loki_amd::hip_check(hipMemcpy(de.get(), ts_e, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::hip_check(hipMemcpy(dv.get(), ts_v, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        ffa->exec.execute_scores(de.as<float>(), dv.as<float>(),
                                 std::span<const size_t>(widths, nwidths), ds.as<float>(),
                                 scores_len, find_max != 0, nullptr);
        loki_amd::hip_check(hipMemcpy(scores, ds.get(), scores_len * 4, hipMemcpyDeviceToHost),
                            "D2H");
    });
}

namespace {

/// Device buffers of one host-side filtered call, copied back sorted by index
/// (the device order depends on scheduling).
struct FilterBuffers {
    loki_amd::DeviceBuffer idx, snr, cnt;
    loki_amd::kern::FilterOut fo{};
    FilterBuffers(float threshold, size_t capacity)
        : idx(std::max<size_t>(capacity, 1) * 4), snr(std::max<size_t>(capacity, 1) * 4),
          cnt(4) {
        fo.threshold = threshold;
        fo.capacity  = static_cast<uint32_t>(capacity);
        fo.count     = cnt.as<uint32_t>();
        fo.indices   = idx.as<uint32_t>();
        fo.snr       = snr.as<float>();
    }
    size_t copy_out(uint32_t* indices, float* scores) const {
        uint32_t n = 0;
        loki_amd::hip_check(hipMemcpy(&n, cnt.get(), 4, hipMemcpyDeviceToHost), "D2H count");
        const size_t k = std::min<size_t>(n, fo.capacity);
        std::vector<uint32_t> vi(k);
        std::vector<float> vs(k);
        if (k > 0) {
            loki_amd::hip_check(hipMemcpy(vi.data(), idx.get(), k * 4, hipMemcpyDeviceToHost),
                                "D2H indices");
            loki_amd::hip_check(hipMemcpy(vs.data(), snr.get(), k * 4, hipMemcpyDeviceToHost),
                                "D2H scores");
        }
        std::vector<size_t> order(k);
        for (size_t i = 0; i < k; ++i) {
            order[i] = i;
        }
        std::sort(order.begin(), order.end(),
                  [&](size_t a, size_t b) { return vi[a] < vi[b]; });
        for (size_t i = 0; i < k; ++i) {
            indices[i] = vi[order[i]];
            scores[i]  = vs[order[i]];
        }
        return n;
    }
};

loki_amd::kern::FilterOut filter_out(float threshold, uint32_t* indices, float* scores,
                                     size_t capacity, uint32_t* count) {
    REQ(count != nullptr, "count is null");
    REQ(capacity == 0 || (indices != nullptr && scores != nullptr), "null output buffer");
    REQ(capacity <= 0xFFFFFFFFULL, "capacity exceeds 32 bits");
    loki_amd::kern::FilterOut fo{};
    fo.threshold = threshold;
    fo.capacity  = static_cast<uint32_t>(capacity);
    fo.count     = count;
    fo.indices   = indices;
    fo.snr       = scores;
    return fo;
}

}  // namespace

int loki_ffa_scores_filter_device(loki_ffa_t ffa, const float* ts_e, const float* ts_v,
                                  const size_t* widths, size_t nwidths, float threshold,
                                  int find_max, uint32_t* indices, float* scores, size_t capacity,
                                  uint32_t* count, void* stream) {
    return guarded([&] {
        REQ(ffa != nullptr && widths != nullptr, "null argument");
        const auto fo = filter_out(threshold, indices, scores, capacity, count);
        ffa->exec.execute_scores(ts_e, ts_v, std::span<const size_t>(widths, nwidths), nullptr,
                                 0, find_max != 0, as_stream(stream), &fo);
    });
}

int loki_ffa_scores_filter_host(loki_ffa_t ffa, const float* ts_e, const float* ts_v,
                                size_t nsamps, const size_t* widths, size_t nwidths,
                                float threshold, int find_max, uint32_t* indices, float* scores,
                                size_t capacity, size_t* count) {
    return guarded([&] {
        REQ(ffa != nullptr && ts_e != nullptr && ts_v != nullptr && widths != nullptr &&
                count != nullptr,
            "null argument");
        REQ(capacity == 0 || (indices != nullptr && scores != nullptr), "null output buffer");
        REQ(capacity <= 0xFFFFFFFFULL, "capacity exceeds 32 bits");
        const auto& pl = ffa->exec.plan();
        REQ(nsamps == pl.config().get_nsamps(), "FFA::execute: ts_e must have size nsamps");
        loki_amd::DeviceGuard g(ffa->exec.device_id());
        loki_amd::DeviceBuffer de(nsamps * 4), dv(nsamps * 4);
        loki_amd::hip_check(hipMemcpy(de.get(), ts_e, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::hip_check(hipMemcpy(dv.get(), ts_v, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        FilterBuffers fb(threshold, capacity);
        ffa->exec.execute_scores(de.as<float>(), dv.as<float>(),
                                 std::span<const size_t>(widths, nwidths), nullptr, 0,
                                 find_max != 0, nullptr, &fb.fo);
        *count = fb.copy_out(indices, scores);
    });
}

int loki_ffa_set_timing(loki_ffa_t ffa, int enable) {
    return guarded([&] {
        REQ(ffa != nullptr, "null ffa");
        ffa->exec.set_timing(enable != 0);
    });
}

int loki_ffa_set_fused(loki_ffa_t ffa, int enable) {
    return guarded([&] {
        REQ(ffa != nullptr, "null ffa");
        ffa->exec.set_fused(enable != 0);
    });
}

int loki_ffa_set_fuse_depth(loki_ffa_t ffa, int depth) {
    return guarded([&] {
        REQ(ffa != nullptr, "null ffa");
        ffa->exec.set_fuse_depth(depth);
    });
}

int loki_ffa_set_two_level_merges(loki_ffa_t ffa, int enable) {
    return guarded([&] {
        REQ(ffa != nullptr, "null ffa");
        ffa->exec.set_two_level_merges(enable != 0);
    });
}

int loki_ffa_set_dft_precision(loki_ffa_t ffa, int mode) {
    return guarded([&] {
        REQ(ffa != nullptr, "null argument");
        REQ(mode == 0 || mode == 1, "dft precision must be 0 (f32) or 1 (split bf16)");
        ffa->exec.set_dft_split_bf16(mode == 1);
    });
}

int loki_ffa_set_exact_norm(loki_ffa_t ffa, int exact) {
    return guarded([&] {
        REQ(ffa != nullptr, "null ffa");
        ffa->exec.set_exact_norm(exact != 0);
    });
}

int loki_ffa_set_fold_storage(loki_ffa_t ffa, int storage) {
    return guarded([&] {
        REQ(ffa != nullptr, "null ffa");
        REQ(storage == LOKI_STORAGE_F32 || storage == LOKI_STORAGE_F16 ||
                storage == LOKI_STORAGE_BF16 || storage == LOKI_STORAGE_BFP16,
            "fold storage must be 0 (f32), 1 (f16), 2 (bf16) or 3 (bfp16)");
        ffa->exec.set_fold_storage(static_cast<loki_amd::FoldStorage>(storage));
    });
}

int loki_ffa_timings(loki_ffa_t ffa, loki_kernel_timing* out, size_t cap, size_t* n) {
    return guarded([&] {
        REQ(ffa != nullptr && n != nullptr, "null argument");
        loki_amd::DeviceGuard g(ffa->exec.device_id());
        const auto t = ffa->exec.timings();
        *n           = t.size();
        REQ(out == nullptr || cap >= t.size(), "output buffer too small");
        if (out != nullptr) {
            for (size_t i = 0; i < t.size(); ++i) {
                out[i] = {static_cast<int>(t[i].kind), t[i].level, t[i].ms};
            }
        }
    });
}

int loki_batch_create(const loki_config_t* cfgs, size_t nseries, const size_t* widths,
                      size_t nwidths, int device_id, loki_batch_t* out) {
    return guarded([&] {
        REQ(cfgs != nullptr && out != nullptr && nseries > 0, "null argument or empty batch");
        REQ(widths != nullptr || nwidths == 0, "null widths");
        std::vector<loki_amd::SearchConfig> v;
        v.reserve(nseries);
        for (size_t i = 0; i < nseries; ++i) {
            REQ(cfgs[i] != nullptr, "null config in batch");
            v.push_back(cfgs[i]->cfg);
        }
        std::vector<loki_amd::SizeType> w(widths, widths + nwidths);
        auto h   = std::make_unique<loki_batch_s>();
        h->batch = std::make_unique<loki_amd::FFABatch>(v, w, device_id);
        for (size_t i = 0; i < nseries; ++i) {
            h->plan_views.push_back(loki_plan_s{h->batch->plan(i)});
        }
        *out = h.release();
    });
}

void loki_batch_destroy(loki_batch_t batch) { delete batch; }

int loki_batch_info(loki_batch_t batch, size_t* nseries, size_t* total_trials,
                    size_t* device_bytes, size_t* launches) {
    return guarded([&] {
        REQ(batch != nullptr, "null batch");
        if (nseries != nullptr) {
            *nseries = batch->batch->size();
        }
        if (total_trials != nullptr) {
            *total_trials = batch->batch->total_trials();
        }
        if (device_bytes != nullptr) {
            *device_bytes = batch->batch->device_bytes();
        }
        if (launches != nullptr) {
            *launches = batch->batch->launches();
        }
    });
}

int loki_batch_plan(loki_batch_t batch, size_t i, loki_plan_t* out_plan_view) {
    return guarded([&] {
        REQ(batch != nullptr && out_plan_view != nullptr, "null argument");
        REQ(i < batch->plan_views.size(), "series index out of range");
        *out_plan_view = &batch->plan_views[i];
    });
}

int loki_batch_series(loki_batch_t batch, size_t i, float** ts_e, float** ts_v, float** scores,
                      size_t* scores_len, size_t* nwidths) {
    return guarded([&] {
        REQ(batch != nullptr, "null batch");
        auto& b = *batch->batch;
        REQ(i < b.size(), "series index out of range");
        if (ts_e != nullptr) {
            *ts_e = b.ts_e(i);
        }
        if (ts_v != nullptr) {
            *ts_v = b.ts_v(i);
        }
        if (scores != nullptr) {
            *scores = b.scores(i);
        }
        if (scores_len != nullptr) {
            *scores_len = b.scores_len(i);
        }
        if (nwidths != nullptr) {
            *nwidths = b.widths(i).size();
        }
    });
}

int loki_batch_set_input_host(loki_batch_t batch, size_t i, const float* ts_e,
                              const float* ts_v, size_t nsamps) {
    return guarded([&] {
        REQ(batch != nullptr && ts_e != nullptr && ts_v != nullptr, "null argument");
        auto& b = *batch->batch;
        REQ(i < b.size(), "series index out of range");
        REQ(nsamps == b.plan(i).config().get_nsamps(), "FFA::execute: ts_e must have size nsamps");
        loki_amd::DeviceGuard g(b.device_id());
        loki_amd::hip_check(hipMemcpy(b.ts_e(i), ts_e, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::hip_check(hipMemcpy(b.ts_v(i), ts_v, nsamps * 4, hipMemcpyHostToDevice), "H2D");
    });
}

int loki_batch_set_input_device(loki_batch_t batch, size_t i, const float* ts_e,
                                const float* ts_v, size_t nsamps, void* stream) {
    return guarded([&] {
        REQ(batch != nullptr && ts_e != nullptr && ts_v != nullptr, "null argument");
        auto& b = *batch->batch;
        REQ(i < b.size(), "series index out of range");
        REQ(nsamps == b.plan(i).config().get_nsamps(), "FFA::execute: ts_e must have size nsamps");
        loki_amd::DeviceGuard g(b.device_id());
        const hipStream_t s = as_stream(stream);
        loki_amd::hip_check(
            hipMemcpyAsync(b.ts_e(i), ts_e, nsamps * 4, hipMemcpyDeviceToDevice, s), "D2D");
        loki_amd::hip_check(
            hipMemcpyAsync(b.ts_v(i), ts_v, nsamps * 4, hipMemcpyDeviceToDevice, s), "D2D");
    });
}

int loki_batch_scores_host(loki_batch_t batch, size_t i, float* out, size_t len) {
    return guarded([&] {
        REQ(batch != nullptr && out != nullptr, "null argument");
        auto& b = *batch->batch;
        REQ(i < b.size(), "series index out of range");
        REQ(len == b.scores_len(i), "scores must have size ncoords * nwidths");
        loki_amd::DeviceGuard g(b.device_id());
        loki_amd::hip_check(hipDeviceSynchronize(), "batch");
        loki_amd::hip_check(hipMemcpy(out, b.scores(i), len * 4, hipMemcpyDeviceToHost), "D2H");
    });
}

int loki_batch_widths(loki_batch_t batch, size_t i, size_t* out, size_t cap, size_t* n) {
    return guarded([&] {
        REQ(batch != nullptr && n != nullptr, "null argument");
        REQ(i < batch->batch->size(), "series index out of range");
        const auto& w = batch->batch->widths(i);
        *n            = w.size();
        REQ(out == nullptr || cap >= w.size(), "output buffer too small");
        if (out != nullptr) {
            std::copy(w.begin(), w.end(), out);
        }
    });
}

int loki_batch_execute(loki_batch_t batch, void* stream) {
    return guarded([&] {
        REQ(batch != nullptr, "null batch");
        batch->batch->execute(as_stream(stream));
    });
}

int loki_batch_set_graph(loki_batch_t batch, int enable) {
    return guarded([&] {
        REQ(batch != nullptr, "null batch");
        batch->batch->set_graph(enable != 0);
    });
}

int loki_batch_set_schedule(loki_batch_t batch, const loki_batch_schedule* sc) {
    return guarded([&] {
        REQ(batch != nullptr && sc != nullptr, "null argument");
        loki_amd::BatchSchedule b;
        b.merge_group = sc->merge_group;
        b.merge_run   = sc->merge_run;
        b.final_group = sc->final_group;
        b.final_run   = sc->final_run;
        b.persistent  = sc->persistent;
        for (int k = 0; k < 5; ++k) {
            b.score_lanes[k] = sc->score_lanes[k];
        }
        batch->batch->set_schedule(b);
    });
}

int loki_batch_errors(loki_batch_t batch, uint32_t* flags) {
    return guarded([&] {
        REQ(batch != nullptr && flags != nullptr, "null argument");
        *flags = batch->batch->errors();
    });
}

int loki_batch_autotune(loki_batch_t batch, void* stream, int repeats,
                        loki_batch_schedule* chosen) {
    return guarded([&] {
        REQ(batch != nullptr, "null batch");
        const auto sc = batch->batch->autotune(as_stream(stream), repeats);
        if (chosen != nullptr) {
            chosen->merge_group = sc.merge_group;
            chosen->merge_run   = sc.merge_run;
            chosen->final_group = sc.final_group;
            chosen->final_run   = sc.final_run;
            chosen->persistent  = sc.persistent;
            for (int k = 0; k < 5; ++k) {
                chosen->score_lanes[k] = sc.score_lanes[k];
            }
        }
    });
}

int loki_generate_ffa_regions(double p_min, double p_max, double tsamp, size_t nbins_min,
                              double eta_min, double octave_scale, size_t nbins_max,
                              loki_ffa_region* out, size_t cap, size_t* n) {
    return guarded([&] {
        REQ(n != nullptr, "null argument");
        const auto r = loki_amd::generate_ffa_regions(p_min, p_max, tsamp, nbins_min, eta_min,
                                                      octave_scale, nbins_max);
        *n = r.size();
        REQ(out == nullptr || cap >= r.size(), "region output too small");
        for (size_t i = 0; out != nullptr && i < r.size(); ++i) {
            out[i] = {r[i].f_start, r[i].f_end, r[i].nbins, r[i].eta};
        }
    });
}

namespace {
void fill_stats(const loki_amd::FFARegionStats& st, loki_region_stats* out) {
    out->max_buffer_size         = st.max_buffer_size();
    out->max_coord_size          = st.max_coord_size();
    out->max_ncoords             = st.max_ncoords();
    out->max_ffa_levels          = st.max_ffa_levels();
    out->max_buffer_size_time    = st.max_buffer_size_time();
    out->max_scores_size         = st.max_scores_size();
    out->write_param_sets_size   = st.write_param_sets_size();
    out->buffer_memory_usage     = st.buffer_memory_usage();
    out->coord_memory_usage      = st.coord_memory_usage();
    out->extra_memory_usage      = st.extra_memory_usage();
    out->freq_sweep_memory_usage = st.freq_sweep_memory_usage();
}
}  // namespace

int loki_region_stats_compute(size_t max_buffer_size, size_t max_coord_size, size_t max_ncoords,
                              size_t max_ffa_levels, size_t n_widths, size_t n_params,
                              size_t n_samps, size_t max_passing_candidates, int use_fourier,
                              int use_gpu, loki_region_stats* out) {
    return guarded([&] {
        REQ(out != nullptr, "null argument");
        fill_stats(loki_amd::FFARegionStats(max_buffer_size, max_coord_size, max_ncoords,
                                            max_ffa_levels, n_widths, n_params, n_samps,
                                            max_passing_candidates, use_fourier != 0,
                                            use_gpu != 0),
                   out);
    });
}

int loki_region_planner_create(loki_config_t cfg, int fold_type, int use_gpu,
                               loki_region_planner_t* out) {
    return guarded([&] {
        REQ(cfg != nullptr && out != nullptr, "null argument");
        *out = new loki_region_planner_s{
            loki_amd::FFARegionPlanner(cfg->cfg, fold_kind(fold_type), use_gpu != 0)};
    });
}

void loki_region_planner_destroy(loki_region_planner_t p) { delete p; }

int loki_region_planner_nregions(loki_region_planner_t p, size_t* n) {
    return guarded([&] {
        REQ(p != nullptr && n != nullptr, "null argument");
        *n = p->planner.nregions();
    });
}

int loki_region_planner_config(loki_region_planner_t p, size_t i, loki_config_t* out) {
    return guarded([&] {
        REQ(p != nullptr && out != nullptr, "null argument");
        REQ(i < p->planner.nregions(), "region index out of range");
        *out = new loki_config_s{p->planner.cfgs()[i]};
    });
}

int loki_region_planner_stats(loki_region_planner_t p, loki_region_stats* out) {
    return guarded([&] {
        REQ(p != nullptr && out != nullptr, "null argument");
        fill_stats(p->planner.stats(), out);
    });
}

int loki_region_planner_chunk(loki_region_planner_t p, size_t i, loki_ffa_chunk_stats* out) {
    return guarded([&] {
        REQ(p != nullptr && out != nullptr, "null argument");
        REQ(i < p->planner.chunk_stats().size(), "chunk index out of range");
        const auto& c = p->planner.chunk_stats()[i];
        *out = {c.nominal_f_start, c.nominal_f_end, c.actual_f_start, c.actual_f_end,
                c.nominal_width,   c.actual_width,  c.total_memory_gb, c.overlap_fraction};
    });
}

int loki_freq_sweep_create(loki_config_t cfg, int device_id, loki_freq_sweep_t* out) {
    return guarded([&] {
        REQ(cfg != nullptr && out != nullptr, "null argument");
        *out = new loki_freq_sweep_s{loki_amd::FFAFreqSweep(cfg->cfg, device_id)};
    });
}

void loki_freq_sweep_destroy(loki_freq_sweep_t s) { delete s; }

int loki_freq_sweep_execute(loki_freq_sweep_t s, const float* ts_e, const float* ts_v,
                            size_t nsamps, size_t* ncandidates) {
    return guarded([&] {
        REQ(s != nullptr && ncandidates != nullptr, "null argument");
        *ncandidates = s->sweep.execute(ts_e, ts_v, nsamps);
    });
}

int loki_freq_sweep_count(loki_freq_sweep_t s, size_t* n) {
    return guarded([&] {
        REQ(s != nullptr && n != nullptr, "null argument");
        *n = s->sweep.snr().size();
    });
}

int loki_freq_sweep_per_region(loki_freq_sweep_t s, size_t* counts, size_t cap) {
    return guarded([&] {
        REQ(s != nullptr && counts != nullptr, "null argument");
        const auto& v = s->sweep.per_region();
        REQ(cap >= s->sweep.planner().nregions(), "per-region output too small");
        for (size_t i = 0; i < s->sweep.planner().nregions(); ++i) {
            counts[i] = i < v.size() ? v[i] : 0;
        }
    });
}

int loki_freq_sweep_results(loki_freq_sweep_t s, float* snr, double* param_sets, size_t cap) {
    return guarded([&] {
        REQ(s != nullptr, "null argument");
        const auto& v = s->sweep.snr();
        const auto& r = s->sweep.rows();
        REQ(cap >= v.size(), "candidate output too small");
        REQ(v.empty() || (snr != nullptr && param_sets != nullptr), "null output buffer");
        std::copy(v.begin(), v.end(), snr);
        std::copy(r.begin(), r.end(), param_sets);
    });
}

int loki_freq_sweep_planner(loki_freq_sweep_t s, loki_region_planner_t* out) {
    return guarded([&] {
        REQ(s != nullptr && out != nullptr, "null argument");
        *out = new loki_region_planner_s{s->sweep.planner()};
    });
}

int loki_brute_fold_time_host(const float* ts_e, const float* ts_v, size_t nsamps,
                              const double* freqs, size_t nfreqs, size_t segment_len,
                              size_t nbins, double tsamp, double t_ref, float* fold,
                              int device_id) {
    return guarded([&] {
        REQ(ts_e != nullptr && ts_v != nullptr && freqs != nullptr && fold != nullptr,
            "null argument");
        REQ(segment_len > 0 && nsamps % segment_len == 0,
            "BruteFold: Number of samples is not a multiple of segment length");
        loki_amd::DeviceGuard g(device_id);
        const size_t out_len = (nsamps / segment_len) * nfreqs * 2 * nbins;
        loki_amd::DeviceBuffer de(nsamps * 4), dv(nsamps * 4), df(out_len * 4);
        loki_amd::hip_check(hipMemcpy(de.get(), ts_e, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::hip_check(hipMemcpy(dv.get(), ts_v, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::brute_fold_time_device(de.as<float>(), dv.as<float>(), nsamps,
                                         std::span<const double>(freqs, nfreqs), segment_len,
                                         nbins, tsamp, t_ref, df.as<float>(), nullptr);
        loki_amd::hip_check(hipMemcpy(fold, df.get(), out_len * 4, hipMemcpyDeviceToHost), "D2H");
    });
}

int loki_brute_fold_time_device(const float* ts_e, const float* ts_v, size_t nsamps,
                                const double* freqs_host, size_t nfreqs, size_t segment_len,
                                size_t nbins, double tsamp, double t_ref, float* fold,
                                void* stream) {
    return guarded([&] {
        REQ(freqs_host != nullptr && fold != nullptr, "null argument");
        loki_amd::brute_fold_time_device(ts_e, ts_v, nsamps,
                                         std::span<const double>(freqs_host, nfreqs),
                                         segment_len, nbins, tsamp, t_ref, fold,
                                         as_stream(stream));
    });
}

int loki_brute_fold_time_device_noalloc(const float* ts_e, const float* ts_v, size_t nsamps,
                                        const double* freqs_host, size_t nfreqs,
                                        size_t segment_len, size_t nbins, double tsamp,
                                        double t_ref, float* fold, void* stream) {
    return guarded([&] {
        REQ(nfreqs < (1ULL << 32) && segment_len < (1ULL << 32),
            "BruteFold: frequency count and segment length must be below 2^32");
        loki_amd::kern::brute_fold_time_noalloc(
            ts_e, ts_v, nsamps, freqs_host, static_cast<uint32_t>(nfreqs),
            static_cast<uint32_t>(segment_len), static_cast<uint32_t>(nbins), tsamp, t_ref, fold,
            as_stream(stream));
    });
}

int loki_brute_fold_fourier_host(const float* ts_e, const float* ts_v, size_t nsamps,
                                 const double* freqs, size_t nfreqs, size_t segment_len,
                                 size_t nbins, double tsamp, double t_ref, float* fold,
                                 int device_id) {
    return guarded([&] {
        REQ(ts_e != nullptr && ts_v != nullptr && freqs != nullptr && fold != nullptr,
            "null argument");
        REQ(segment_len > 0 && nsamps % segment_len == 0,
            "BruteFold: Number of samples is not a multiple of segment length");
        loki_amd::DeviceGuard g(device_id);
        const size_t out_len = (nsamps / segment_len) * nfreqs * 2 * (nbins / 2 + 1);
        loki_amd::DeviceBuffer de(nsamps * 4), dv(nsamps * 4), df(out_len * 8);
        loki_amd::hip_check(hipMemcpy(de.get(), ts_e, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::hip_check(hipMemcpy(dv.get(), ts_v, nsamps * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::brute_fold_fourier_device(de.as<float>(), dv.as<float>(), nsamps,
                                            std::span<const double>(freqs, nfreqs), segment_len,
                                            nbins, tsamp, t_ref, df.as<float2>(), nullptr);
        loki_amd::hip_check(hipMemcpy(fold, df.get(), out_len * 8, hipMemcpyDeviceToHost), "D2H");
    });
}

int loki_brute_fold_fourier_device(const float* ts_e, const float* ts_v, size_t nsamps,
                                   const double* freqs_host, size_t nfreqs, size_t segment_len,
                                   size_t nbins, double tsamp, double t_ref, float* fold,
                                   void* stream) {
    return guarded([&] {
        REQ(freqs_host != nullptr && fold != nullptr, "null argument");
        loki_amd::brute_fold_fourier_device(ts_e, ts_v, nsamps,
                                            std::span<const double>(freqs_host, nfreqs),
                                            segment_len, nbins, tsamp, t_ref,
                                            reinterpret_cast<float2*>(fold), as_stream(stream));
    });
}

int loki_irfft_batch_device(const float* complex_in, float* real_out, size_t batch_size,
                            size_t n_real, void* stream) {
    return guarded([&] {
        REQ(complex_in != nullptr && real_out != nullptr, "null argument");
        REQ(n_real >= 1 && n_real <= 4096, "IRFFT: n_real must be in [1, 4096]");
        loki_amd::kern::irfft_rows(reinterpret_cast<const float2*>(complex_in), real_out,
                                   batch_size, static_cast<uint32_t>(n_real), as_stream(stream));
        loki_amd::hip_check(hipGetLastError(), "irfft launch");
    });
}

int loki_rfft_batch_device(const float* real_in, float* complex_out, size_t batch_size,
                           size_t n_real, void* stream) {
    return guarded([&] {
        REQ(real_in != nullptr && complex_out != nullptr, "null argument");
        REQ(n_real >= 1 && n_real <= 4096, "RFFT: n_real must be in [1, 4096]");
        loki_amd::kern::rfft_rows(real_in, reinterpret_cast<float2*>(complex_out), batch_size,
                                  static_cast<uint32_t>(n_real), as_stream(stream));
        loki_amd::hip_check(hipGetLastError(), "rfft launch");
    });
}

size_t loki_refine_workspace_bytes(size_t nsamps, size_t ncand, size_t nbins) {
    return loki_amd::kern::refine_workspace_bytes(nsamps, ncand, static_cast<uint32_t>(nbins));
}

int loki_refine_score_device(const float* ts_e, const float* ts_v, size_t nsamps, double tsamp,
                             double t_ref, const double* theta, size_t ncand, size_t nbins,
                             size_t width, float* scores, float* adj_state, void* workspace,
                             size_t workspace_bytes, void* stream) {
    return guarded([&] {
        REQ(ncand >= 1 && ncand < 65536, "refine: 1 <= candidates < 65536");
        REQ(ts_e != nullptr && ts_v != nullptr && theta != nullptr && scores != nullptr &&
                workspace != nullptr,
            "null argument");
        REQ(nbins <= 128, "refine: nbins must be 16, 32, 64 or 128");
        REQ(workspace_bytes >= loki_amd::kern::refine_workspace_bytes(
                                   nsamps, ncand, static_cast<uint32_t>(nbins)),
            "refine: workspace_bytes too small");
        REQ(tsamp > 0.0, "refine: tsamp must be positive");
        loki_amd::kern::refine_forward(ts_e, ts_v, nsamps, tsamp, t_ref, theta, ncand,
                                       static_cast<uint32_t>(nbins),
                                       static_cast<uint32_t>(width < 1024 ? width : 1024), scores,
                                       reinterpret_cast<float2*>(adj_state), workspace,
                                       as_stream(stream));
        loki_amd::hip_check(hipGetLastError(), "refine score launch");
    });
}

int loki_refine_grad_device(const float* ts_e, const float* ts_v, size_t nsamps, double tsamp,
                            double t_ref, const double* theta, size_t ncand, size_t nbins,
                            const float* adj_state, double* grad, void* workspace,
                            size_t workspace_bytes, void* stream) {
    return guarded([&] {
        REQ(ncand >= 1 && ncand < 65536, "refine: 1 <= candidates < 65536");
        REQ(ts_e != nullptr && ts_v != nullptr && theta != nullptr && adj_state != nullptr &&
                grad != nullptr && workspace != nullptr,
            "null argument");
        REQ(nbins <= 128, "refine: nbins must be 16, 32, 64 or 128");
        REQ(nsamps >= 1, "refine: empty series");
        REQ(workspace_bytes >= loki_amd::kern::refine_workspace_bytes(
                                   nsamps, ncand, static_cast<uint32_t>(nbins)),
            "refine: workspace_bytes too small");
        REQ(tsamp > 0.0, "refine: tsamp must be positive");
        loki_amd::kern::refine_adjoint(ts_e, ts_v, nsamps, tsamp, t_ref, theta, ncand,
                                       static_cast<uint32_t>(nbins),
                                       reinterpret_cast<const float2*>(adj_state), grad,
                                       workspace, as_stream(stream));
        loki_amd::hip_check(hipGetLastError(), "refine grad launch");
    });
}

int loki_snr_boxcar_3d_device(const float* folds, size_t nprof, size_t nbins,
                              const size_t* widths, size_t nwidths, float* scores, int find_max,
                              void* stream) {
    return guarded([&] {
        REQ(folds != nullptr && scores != nullptr && widths != nullptr, "null argument");
        const auto wt = loki_amd::kern::make_width_table(widths, nwidths, nbins);
        loki_amd::kern::snr_boxcar_3d(folds, nprof, static_cast<uint32_t>(nbins), wt, scores,
                                      find_max != 0, as_stream(stream));
        loki_amd::hip_check(hipGetLastError(), "score launch");
    });
}

int loki_snr_boxcar_3d_host(const float* folds, size_t nprof, size_t nbins,
                            const size_t* widths, size_t nwidths, float* scores, int find_max,
                            int device_id) {
    return guarded([&] {
        REQ(folds != nullptr && scores != nullptr && widths != nullptr, "null argument");
        const auto wt = loki_amd::kern::make_width_table(widths, nwidths, nbins);
        loki_amd::DeviceGuard g(device_id);
        const size_t in_len  = nprof * 2 * nbins;
        const size_t out_len = find_max ? nprof : nprof * nwidths;
        loki_amd::DeviceBuffer di(in_len * 4), dout(out_len * 4);
        loki_amd::hip_check(hipMemcpy(di.get(), folds, in_len * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::kern::snr_boxcar_3d(di.as<float>(), nprof, static_cast<uint32_t>(nbins), wt,
                                      dout.as<float>(), find_max != 0, nullptr);
        loki_amd::hip_check(hipGetLastError(), "score launch");
        loki_amd::hip_check(hipMemcpy(scores, dout.get(), out_len * 4, hipMemcpyDeviceToHost),
                            "D2H");
    });
}

int loki_selftest_exact_norm_sweep(int device_id, uint32_t bits_lo, uint32_t bits_hi,
                                   uint64_t* counts, uint32_t* first_bad) {
    return guarded([&] {
        REQ(counts != nullptr && first_bad != nullptr, "null argument");
        loki_amd::DeviceGuard g(device_id);
        loki_amd::DeviceBuffer dc(4 * sizeof(uint64_t)), df(2 * sizeof(uint32_t));
        const uint64_t zero[4]  = {0, 0, 0, 0};
        const uint32_t none[2]  = {0xFFFFFFFFU, 0xFFFFFFFFU};
        loki_amd::hip_check(hipMemcpy(dc.get(), zero, sizeof(zero), hipMemcpyHostToDevice), "H2D");
        loki_amd::hip_check(hipMemcpy(df.get(), none, sizeof(none), hipMemcpyHostToDevice), "H2D");
        loki_amd::kern::selftest_exact_norm_sweep(bits_lo, bits_hi,
                                                  dc.as<unsigned long long>(),
                                                  df.as<uint32_t>(), nullptr);
        loki_amd::hip_check(hipGetLastError(), "self-test launch");
        loki_amd::hip_check(hipMemcpy(counts, dc.get(), sizeof(zero), hipMemcpyDeviceToHost),
                            "D2H");
        loki_amd::hip_check(hipMemcpy(first_bad, df.get(), sizeof(none), hipMemcpyDeviceToHost),
                            "D2H");
    });
}

int loki_selftest_exact_norm_values(int device_id, const uint32_t* bits, size_t n, float* out) {
    return guarded([&] {
        REQ(bits != nullptr && out != nullptr, "null argument");
        loki_amd::DeviceGuard g(device_id);
        if (n == 0) {
            return;
        }
        loki_amd::DeviceBuffer db(n * 4), dout(n * 16);
        loki_amd::hip_check(hipMemcpy(db.get(), bits, n * 4, hipMemcpyHostToDevice), "H2D");
        loki_amd::kern::selftest_exact_norm_values(db.as<uint32_t>(), n, dout.as<float>(),
                                                   nullptr);
        loki_amd::hip_check(hipGetLastError(), "self-test launch");
        loki_amd::hip_check(hipMemcpy(out, dout.get(), n * 16, hipMemcpyDeviceToHost), "D2H");
    });
}

int loki_snr_boxcar_3d_filter_device(const float* folds, size_t nprof, size_t nbins,
                                     const size_t* widths, size_t nwidths, float threshold,
                                     int find_max, uint32_t* indices, float* scores,
                                     size_t capacity, uint32_t* count, void* stream) {
    return guarded([&] {
        REQ(folds != nullptr && widths != nullptr, "null argument");
        const auto wt = loki_amd::kern::make_width_table(widths, nwidths, nbins);
        const auto fo = filter_out(threshold, indices, scores, capacity, count);
        loki_amd::kern::zero_counter(count, as_stream(stream));
        loki_amd::kern::snr_boxcar_3d(folds, nprof, static_cast<uint32_t>(nbins), wt, nullptr,
                                      find_max != 0, as_stream(stream), &fo);
        loki_amd::hip_check(hipGetLastError(), "score filter launch");
    });
}

namespace {
void check_seed(size_t nbins, size_t nleaves) {
    REQ(nbins >= 1 && nbins <= 0xFFFFFFFFULL, "nbins must be in [1, 2^32)");
    REQ(nleaves < (1ULL << 32), "n_leaves must be < 2^32");
}
}  // namespace

int loki_shift_add_linear_batch_device(const float* folds_tree, const uint32_t* indices_tree,
                                       const uint8_t* validation_mask, const float* folds_ffa,
                                       const uint32_t* indices_ffa, const float* phase_shift,
                                       float* folds_out, size_t nbins, size_t n_leaves,
                                       uint32_t physical_start_idx, uint32_t capacity,
                                       size_t ffa_rows, void* stream) {
    return guarded([&] {
        check_seed(nbins, n_leaves);
        REQ(n_leaves == 0 || (folds_tree && indices_tree && folds_ffa && indices_ffa &&
                              phase_shift && folds_out),
            "null argument");
        loki_amd::kern::seed_linear(folds_tree, indices_tree, validation_mask, folds_ffa,
                                    indices_ffa, phase_shift, folds_out,
                                    static_cast<uint32_t>(nbins), n_leaves, physical_start_idx,
                                    capacity, ffa_rows, as_stream(stream));
        loki_amd::hip_check(hipGetLastError(), "shift_add_linear launch");
    });
}

int loki_shift_add_linear_complex_batch_device(
    const void* folds_tree, const uint32_t* indices_tree, const uint8_t* validation_mask,
    const void* folds_ffa, const uint32_t* indices_ffa, const float* phase_shift,
    void* folds_out, size_t nbins_f, size_t nbins, size_t n_leaves, uint32_t physical_start_idx,
    uint32_t capacity, size_t ffa_rows, void* stream) {
    return guarded([&] {
        check_seed(nbins, n_leaves);
        REQ(nbins_f == nbins / 2 + 1, "nbins_f must be nbins / 2 + 1");
        REQ(n_leaves == 0 || (folds_tree && indices_tree && folds_ffa && indices_ffa &&
                              phase_shift && folds_out),
            "null argument");
        loki_amd::kern::seed_linear_complex(
            static_cast<const float2*>(folds_tree), indices_tree, validation_mask,
            static_cast<const float2*>(folds_ffa), indices_ffa, phase_shift,
            static_cast<float2*>(folds_out), static_cast<uint32_t>(nbins_f),
            static_cast<uint32_t>(nbins), n_leaves, physical_start_idx, capacity, ffa_rows,
            as_stream(stream));
        loki_amd::hip_check(hipGetLastError(), "shift_add_linear_complex launch");
    });
}

int loki_shift_add_ascend_linear_batch_device(const float* folds_ffa,
                                              const uint32_t* indices_segment,
                                              const uint32_t* indices_ffa,
                                              const float* phase_shift, float* folds_tree,
                                              size_t nbins, size_t n_coords_init,
                                              size_t n_leaves, size_t n_segments,
                                              size_t ffa_segments, void* stream) {
    return guarded([&] {
        check_seed(nbins, n_leaves);
        REQ(n_segments <= 0xFFFFFFFFULL, "n_segments must be < 2^32");
        REQ(n_leaves == 0 || (folds_tree && (n_segments == 0 || (folds_ffa && indices_segment &&
                                                                 indices_ffa && phase_shift))),
            "null argument");
        loki_amd::kern::seed_ascend(folds_ffa, indices_segment, indices_ffa, phase_shift,
                                    folds_tree, static_cast<uint32_t>(nbins), n_coords_init,
                                    n_leaves, static_cast<uint32_t>(n_segments), ffa_segments,
                                    as_stream(stream));
        loki_amd::hip_check(hipGetLastError(), "shift_add_ascend_linear launch");
    });
}

int loki_shift_add_ascend_linear_complex_batch_device(
    const void* folds_ffa, const uint32_t* indices_segment, const uint32_t* indices_ffa,
    const float* phase_shift, void* folds_tree, size_t nbins_f, size_t nbins,
    size_t n_coords_init, size_t n_leaves, size_t n_segments, size_t ffa_segments,
    void* stream) {
    return guarded([&] {
        check_seed(nbins, n_leaves);
        REQ(nbins_f == nbins / 2 + 1, "nbins_f must be nbins / 2 + 1");
        REQ(n_segments <= 0xFFFFFFFFULL, "n_segments must be < 2^32");
        REQ(n_leaves == 0 || (folds_tree && (n_segments == 0 || (folds_ffa && indices_segment &&
                                                                 indices_ffa && phase_shift))),
            "null argument");
        loki_amd::kern::seed_ascend_complex(
            static_cast<const float2*>(folds_ffa), indices_segment, indices_ffa, phase_shift,
            static_cast<float2*>(folds_tree), static_cast<uint32_t>(nbins_f),
            static_cast<uint32_t>(nbins), n_coords_init, n_leaves,
            static_cast<uint32_t>(n_segments), ffa_segments, as_stream(stream));
        loki_amd::hip_check(hipGetLastError(), "shift_add_ascend_linear_complex launch");
    });
}

int loki_snr_boxcar_3d_filter_host(const float* folds, size_t nprof, size_t nbins,
                                   const size_t* widths, size_t nwidths, float threshold,
                                   int find_max, uint32_t* indices, float* scores,
                                   size_t capacity, size_t* count, int device_id) {
    return guarded([&] {
        REQ(folds != nullptr && widths != nullptr && count != nullptr, "null argument");
        REQ(capacity == 0 || (indices != nullptr && scores != nullptr), "null output buffer");
        REQ(capacity <= 0xFFFFFFFFULL, "capacity exceeds 32 bits");
        const auto wt = loki_amd::kern::make_width_table(widths, nwidths, nbins);
        loki_amd::DeviceGuard g(device_id);
        const size_t in_len = nprof * 2 * nbins;
        loki_amd::DeviceBuffer di(std::max<size_t>(in_len, 1) * 4);
        if (in_len > 0) {
            loki_amd::hip_check(hipMemcpy(di.get(), folds, in_len * 4, hipMemcpyHostToDevice),
                                "H2D");
        }
        FilterBuffers fb(threshold, capacity);
        loki_amd::kern::zero_counter(fb.fo.count, nullptr);
        loki_amd::kern::snr_boxcar_3d(di.as<float>(), nprof, static_cast<uint32_t>(nbins), wt,
                                      nullptr, find_max != 0, nullptr, &fb.fo);
        loki_amd::hip_check(hipGetLastError(), "score filter launch");
        *count = fb.copy_out(indices, scores);
    });
}

int loki_snr_boxcar_2d_max_host(const float* folds, size_t nprof, size_t nbins,
                                const size_t* widths, size_t nwidths, float stdnoise,
                                float* scores, int device_id) {
    return guarded([&] {
        REQ(folds != nullptr && scores != nullptr && widths != nullptr, "null argument");
        const auto wt = loki_amd::kern::make_width_table(widths, nwidths, nbins);
        loki_amd::DeviceGuard g(device_id);
        loki_amd::DeviceBuffer di(nprof * nbins * 4), dout(nprof * 4);
        loki_amd::hip_check(hipMemcpy(di.get(), folds, nprof * nbins * 4, hipMemcpyHostToDevice),
                            "H2D");
        loki_amd::kern::snr_boxcar_2d_max(di.as<float>(), nprof, static_cast<uint32_t>(nbins),
                                          wt, stdnoise, dout.as<float>(), nullptr);
        loki_amd::hip_check(hipGetLastError(), "score launch");
        loki_amd::hip_check(hipMemcpy(scores, dout.get(), nprof * 4, hipMemcpyDeviceToHost),
                            "D2H");
    });
}

}  // extern "C"
