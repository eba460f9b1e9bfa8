// Batched multi-series FFA: arenas, task tables, grouped launches, graph
// capture and the schedule autotuner.
#include "batch.hpp"

#include <algorithm>
#include <cstring>
#include <exception>
#include <map>
#include <string>

namespace loki_amd {
namespace {

template <typename T>
SizeType append(std::vector<uint8_t>& blob, const std::vector<T>& v) {
    const SizeType off = (blob.size() + 255) / 256 * 256;  // 256-byte aligned tables
    blob.resize(off + v.size() * sizeof(T));
    if (!v.empty()) {
        std::memcpy(blob.data() + off, v.data(), v.size() * sizeof(T));
    }
    return off;
}

SizeType ceil_div(SizeType a, SizeType b) { return (a + b - 1) / b; }

}  // namespace

FFABatch::FFABatch(const std::vector<SearchConfig>& cfgs, const std::vector<SizeType>& widths,
                   int device_id)
    : m_device(device_id) {
    LOKI_REQUIRE(!cfgs.empty(), "FFABatch: no series");
    DeviceGuard g(device_id);
    SizeType in_total = 0, fold_total = 0, coord_total = 0, score_total = 0;
    m_series.reserve(cfgs.size());
    for (const auto& cfg : cfgs) {
        Series sr{FFAPlan(cfg, FoldKind::kTime), {}, {}, 0, {0, 0}, 0, 0, 0};
        const auto& pl = sr.plan;
        const SizeType nb = cfg.get_nbins();
        LOKI_REQUIRE(cfg.get_nparams() == 1, "FFABatch: frequency-only searches only");
        LOKI_REQUIRE(nb == 8 || nb == 16 || nb == 32 || nb == 64 || nb == 128,
                     "FFABatch: nbins must be 8, 16, 32, 64 or 128");
        sr.nlevels = pl.n_levels();
        LOKI_REQUIRE(sr.nlevels >= 2, "FFABatch: levels must be greater than or equal to 2");
        LOKI_REQUIRE(pl.nsegments().back() == 1, "compute_ffa_scores: nsegments must be 1 for scores");
        if (widths.empty()) {
            sr.widths = cfg.get_scoring_widths();
        } else {
            for (SizeType w : widths) {
                if (w >= 1 && w < nb) {
                    sr.widths.push_back(w);
                }
            }
        }
        LOKI_REQUIRE(!sr.widths.empty(), "FFABatch: no boxcar width below nbins");
        for (SizeType l = 1; l < sr.nlevels; ++l) {
            LOKI_REQUIRE(pl.ncoords()[l - 1] < (1ULL << 32), "FFABatch: grid exceeds 2^32");
        }
        SizeType buf = 0;
        for (SizeType l = 0; l + 1 < sr.nlevels; ++l) {
            buf = std::max(buf, pl.level_size(l));
        }
        sr.in_off      = in_total;
        in_total      += 2 * ((cfg.get_nsamps() + 63) / 64 * 64);
        sr.buf_off[0]  = fold_total;
        sr.buf_off[1]  = fold_total + (buf + 63) / 64 * 64;
        fold_total    += 2 * ((buf + 63) / 64 * 64);
        sr.coord_off   = coord_total;
        coord_total   += pl.coord_size();
        sr.score_off   = score_total;
        score_total   += (pl.ncoords().back() * sr.widths.size() + 63) / 64 * 64;
        m_series.push_back(std::move(sr));
    }
    m_in.reserve(in_total * sizeof(float));
    m_fold.reserve(fold_total * sizeof(float));
    m_coords.reserve(std::max<SizeType>(coord_total, 1) * sizeof(FFACoordFreq));
    m_scores.reserve(score_total * sizeof(float));
    // brute-fold bucket lists (FP64 phases), once per series
    hipStream_t s = nullptr;
    for (auto& sr : m_series) {
        const auto freqs   = sr.plan.compute_param_grid(0).back();
        const auto& cfg    = sr.plan.config();
        const SizeType nfr = freqs.size();
        LOKI_REQUIRE(sr.plan.segment_lens()[0] < (1ULL << 32) && nfr < (1ULL << 32),
                     "FFABatch: brute fold too large");
        m_freqs.reserve(nfr * sizeof(double));
        hip_check(hipMemcpy(m_freqs.get(), freqs.data(), nfr * sizeof(double),
                            hipMemcpyHostToDevice),
                  "copy brute-fold frequencies");
        sr.lists = kern::build_brute_fold_lists(
            m_freqs.as<double>(), static_cast<uint32_t>(nfr),
            static_cast<uint32_t>(sr.plan.segment_lens()[0]), static_cast<uint32_t>(cfg.get_nbins()),
            cfg.get_tsamp(), 0.0, s);
    }
    build_tasks();
}

FFABatch::~FFABatch() { drop_graph(); }

float* FFABatch::ts_e(SizeType i) const { return m_in.as<float>() + m_series.at(i).in_off; }

float* FFABatch::ts_v(SizeType i) const {
    const auto& sr = m_series.at(i);
    return m_in.as<float>() + sr.in_off + (sr.plan.config().get_nsamps() + 63) / 64 * 64;
}

float* FFABatch::scores(SizeType i) const {
    return m_scores.as<float>() + m_series.at(i).score_off;
}

SizeType FFABatch::scores_len(SizeType i) const {
    const auto& sr = m_series.at(i);
    return sr.plan.ncoords().back() * sr.widths.size();
}

SizeType FFABatch::total_trials() const noexcept {
    SizeType n = 0;
    for (const auto& sr : m_series) {
        n += sr.plan.ncoords().back();
    }
    return n;
}

SizeType FFABatch::device_bytes() const noexcept {
    return m_in.bytes() + m_fold.bytes() + m_coords.bytes() + m_scores.bytes() +
           m_tasks.bytes();
}

void FFABatch::set_schedule(const BatchSchedule& sc) {
    LOKI_REQUIRE(sc.merge_group >= 1 && sc.merge_run >= 1 && sc.final_group >= 1 &&
                     sc.final_run >= 1,
                 "FFABatch: schedule knobs must be positive");
    LOKI_REQUIRE(sc.persistent == 0,
                 "FFABatch: the persistent launch is disabled (its cross-workgroup completion "
                 "counters stalled on MI355X); use the grouped launches (persistent = 0)");
    for (uint32_t k = 0; k < 5; ++k) {
        const uint32_t lanes = sc.score_lanes[k];
        LOKI_REQUIRE(lanes == 0 || kern::final_wide_supported(8U << k, lanes),
                     "FFABatch: score_lanes must be 0, 1, 2 or 4, with nbins >= 4 * lanes and "
                     "nbins <= 32 * lanes above 32 bins");
    }
    DeviceGuard g(m_device);
    const BatchSchedule prev = m_sched;
    m_sched                  = sc;
    try {
        build_tasks();
    } catch (...) {
        // a schedule the batch cannot run (e.g. a last level over the LDS budget): keep the
        // previous one in force
        m_sched = prev;
        build_tasks();
        throw;
    }
}

void FFABatch::set_graph(bool on) {
    m_use_graph = on;
    if (!on) {
        drop_graph();
    }
}

void FFABatch::set_spin_limit(uint32_t polls) {
    m_spin_limit = polls;
    m_tree.spin_limit = polls;
    drop_graph();  // the launch arguments are baked into a captured graph
}

uint32_t FFABatch::errors() {
    DeviceGuard g(m_device);
    if (m_sync.get() == nullptr) {
        return 0;
    }
    hip_check(hipDeviceSynchronize(), "batch: wait for the device");
    uint32_t flag = 0;
    hip_check(hipMemcpy(&flag, m_sync.as<uint32_t>() + 2, sizeof(flag), hipMemcpyDeviceToHost),
              "batch: read the error flag");
    if (flag != 0) {
        const uint32_t zero = 0;
        hip_check(hipMemcpy(m_sync.as<uint32_t>() + 2, &zero, sizeof(zero),
                            hipMemcpyHostToDevice),
                  "batch: clear the error flag");
    }
    return flag;
}

void FFABatch::drop_graph() {
    if (m_exec != nullptr) {
        (void)hipGraphExecDestroy(m_exec);
        m_exec = nullptr;
    }
    if (m_graph != nullptr) {
        (void)hipGraphDestroy(m_graph);
        m_graph = nullptr;
    }
}

void FFABatch::build_tasks() {
    drop_graph();
    m_launches.clear();
    std::vector<uint8_t> blob;
    if (m_sched.persistent != 0) {
        std::vector<SizeType> offs;
        build_tree(blob, offs);
        // a previous execute() on any stream may still read the old tables
        hip_check(hipDeviceSynchronize(), "batch: wait before rewriting the task tables");
        m_tasks.reserve(std::max<SizeType>(blob.size(), 1));
        hip_check(hipMemcpy(m_tasks.get(), blob.data(), blob.size(), hipMemcpyHostToDevice),
                  "copy batch task tables");
        const auto* base  = static_cast<const uint8_t*>(m_tasks.get());
        m_tree.items      = reinterpret_cast<const kern::TreeItem*>(base + offs[0]);
        m_tree.resolve    = reinterpret_cast<const kern::ResolveTask*>(base + offs[1]);
        m_tree.fold       = reinterpret_cast<const kern::FoldTask*>(base + offs[2]);
        m_tree.merge      = reinterpret_cast<const kern::MergeTask*>(base + offs[3]);
        m_tree.final_     = reinterpret_cast<const kern::FinalTask*>(base + offs[4]);
        const SizeType nsync = kern::kTreeSyncHeader + m_tree.nitems;
        m_sync.reserve(nsync * sizeof(uint32_t));
        hip_check(hipMemset(m_sync.get(), 0, nsync * sizeof(uint32_t)), "zero batch counters");
        m_tree.sync       = m_sync.as<uint32_t>();
        m_tree.spin_limit = m_spin_limit;
        m_tree_blocks     = std::min(m_tree.ntickets, kern::tree_resident_blocks(m_tree_lds));
        hip_check(hipDeviceSynchronize(), "batch: task tables");
        return;
    }
    float* fold        = m_fold.as<float>();
    auto* coords       = m_coords.as<FFACoordFreq>();
    auto buf           = [&](const Series& sr, SizeType level) {
        return fold + sr.buf_off[level % 2];
    };

    // 1. shift tables of levels 1..L-2 (the last level resolves in-kernel)
    {
        std::vector<kern::ResolveTask> tasks;
        uint32_t block0 = 0;
        for (const auto& sr : m_series) {
            for (SizeType l = 1; l + 1 < sr.nlevels; ++l) {
                const SizeType n = sr.plan.ncoords()[l];
                tasks.push_back({resolve_level(sr.plan, l),
                                 coords + sr.coord_off + sr.plan.ncoords_offsets64()[l], n,
                                 block0});
                block0 += static_cast<uint32_t>(ceil_div(n, 256));
            }
        }
        if (!tasks.empty()) {
            m_launches.push_back({Stage::kResolve, 0, static_cast<uint32_t>(tasks.size()),
                                  block0, append(blob, tasks), 0, 0, 0});
        }
    }
    // 2. brute folds: list-path short segments, and run-path long segments by
    //    pairs per lane (one grouped launch per kind)
    for (uint32_t kind : {0U, 1U}) {  // 0: list path, 1: run path
        std::vector<kern::FoldTask> tasks;
        uint64_t block0 = 0;
        uint32_t max_nbins = 0;
        for (const auto& sr : m_series) {
            const auto& L = sr.lists;
            if ((L.use_runs ? 1U : 0U) != kind) {
                continue;
            }
            const uint32_t bps = kern::brute_fold_blocks_per_segment(L);
            kern::FoldTask t{};
            t.ts_e   = ts_e(static_cast<SizeType>(&sr - m_series.data()));
            t.ts_v   = ts_v(static_cast<SizeType>(&sr - m_series.data()));
            t.list   = L.list.as<uint32_t>();
            t.base   = L.base.as<uint64_t>();
            t.counts = L.counts.as<uint32_t>();
            t.freqs  = L.freqs.as<double>();
            t.tsamp  = L.tsamp;
            t.t_ref  = L.t_ref;
            t.fold   = buf(sr, 0);
            t.scale  = 1.0F;
            t.seglen = L.seglen;
            t.nfreqs = L.nfreqs;
            t.nbins  = L.nbins;
            t.bps    = bps;
            t.block0 = static_cast<uint32_t>(block0);
            tasks.push_back(t);
            block0 += sr.plan.nsegments()[0] * bps;
            max_nbins = std::max(max_nbins, L.nbins);
        }
        LOKI_REQUIRE(block0 < (1ULL << 31), "FFABatch: brute-fold grid too large");
        if (!tasks.empty()) {
            // fold launches keep the largest nbins of their tasks in `stride`
            m_launches.push_back({Stage::kFold, kind, static_cast<uint32_t>(tasks.size()),
                                  static_cast<uint32_t>(block0), append(blob, tasks),
                                  max_nbins, 0, 0});
        }
    }
    // 3. merges, one launch per level index and nbins class
    SizeType max_levels = 0;
    std::map<uint32_t, int> classes;
    for (const auto& sr : m_series) {
        max_levels = std::max(max_levels, sr.nlevels);
        classes[static_cast<uint32_t>(sr.plan.config().get_nbins())] = 1;
    }
    for (SizeType k = 1; k + 1 < max_levels; ++k) {
        for (const auto& [nb, unused] : classes) {
            (void)unused;
            std::vector<kern::MergeTask> tasks;
            uint64_t block0 = 0;
            for (const auto& sr : m_series) {
                if (sr.plan.config().get_nbins() != nb || k + 1 >= sr.nlevels) {
                    continue;
                }
                const auto& pl = sr.plan;
                kern::MergeTask t{};
                t.in          = buf(sr, k - 1);
                t.out         = buf(sr, k);
                t.coords      = nullptr;
                t.coords_freq = coords + sr.coord_off + pl.ncoords_offsets64()[k];
                t.ncur        = pl.ncoords()[k];
                t.nprev       = pl.ncoords()[k - 1];
                t.ts = kern::make_schedule(t.ncur, pl.param_counts()[k].back(), pl.nsegments()[k],
                                           kern::merge_tile(nb), m_sched.merge_group,
                                           m_sched.merge_run);
                t.scale  = 1.0F;
                t.block0 = static_cast<uint32_t>(block0);
                tasks.push_back(t);
                block0 += t.ts.nblocks;
            }
            LOKI_REQUIRE(block0 < (1ULL << 31), "FFABatch: merge grid too large");
            if (!tasks.empty()) {
                m_launches.push_back({Stage::kMerge, nb, static_cast<uint32_t>(tasks.size()),
                                      static_cast<uint32_t>(block0), append(blob, tasks), 0, 0, 0});
            }
        }
    }
    // 4. fused last levels (shift tables + merge + S/N), one launch per nbins
    //    class, with the class's boxcar-width schedule (lanes per trial)
    for (const auto& [nb, unused] : classes) {
        (void)unused;
        const uint32_t lanes = m_sched.score_lanes[__builtin_ctz(nb) - 3];
        const uint32_t rows  = lanes == 0 ? kern::kFinalRows : kern::final_wide_rows(lanes);
        std::vector<kern::FinalTask> tasks;
        uint64_t block0 = 0;
        uint32_t stride = 1;
        for (const auto& sr : m_series) {
            if (sr.plan.config().get_nbins() == nb) {
                const auto wt = kern::make_width_table(sr.widths.data(), sr.widths.size(), nb);
                stride        = std::max(stride, kern::final_lds_stride(nb, wt.wmax));
            }
        }
        for (const auto& sr : m_series) {
            if (sr.plan.config().get_nbins() != nb) {
                continue;
            }
            const auto& pl  = sr.plan;
            const SizeType L = sr.nlevels;
            kern::FinalTask t{};
            t.in     = buf(sr, L - 2);
            t.scores = m_scores.as<float>() + sr.score_off;
            t.nprev  = pl.ncoords()[L - 2];
            t.lv     = resolve_level(pl, L - 1);
            for (int q = 0; q < 3; ++q) {
                t.fg.dv[q] = (t.lv.hi[q] - t.lv.lo[q]) / static_cast<double>(t.lv.n_cur[q]);
            }
            t.fg.stride = stride;
            t.ts = kern::make_schedule(pl.ncoords().back(), t.lv.n_cur[2], 1, rows,
                                       m_sched.final_group, m_sched.final_run);
            t.wt      = kern::make_width_table(sr.widths.data(), sr.widths.size(), nb);
            t.unscale = 1.0F;
            t.block0  = static_cast<uint32_t>(block0);
            tasks.push_back(t);
            block0 += t.ts.nblocks;
        }
        LOKI_REQUIRE(block0 < (1ULL << 31), "FFABatch: last-level grid too large");
        const SizeType lds = lanes == 0 ? 0 : kern::final_wide_lds_bytes(lanes, stride);
        LOKI_REQUIRE(lds <= 64 * 1024,
                     "FFABatch: a last-level workgroup needs more than 64 KiB of LDS (more "
                     "score_lanes for the widest rows)");
        if (!tasks.empty()) {
            m_launches.push_back({Stage::kFinal, nb, static_cast<uint32_t>(tasks.size()),
                                  static_cast<uint32_t>(block0), append(blob, tasks), stride,
                                  lanes, lds});
        }
    }
    // a previous execute() on any stream may still read the old task tables:
    // let the device finish before they are replaced
    hip_check(hipDeviceSynchronize(), "batch: wait before rewriting the task tables");
    m_tasks.reserve(std::max<SizeType>(blob.size(), 1));
    hip_check(hipMemcpy(m_tasks.get(), blob.data(), blob.size(), hipMemcpyHostToDevice),
              "copy batch task tables");
}

// Items of the persistent launch, in ticket order: every shift table, every
// brute fold (largest first, so the tail of the fold stage is short), then
// stage by stage the merge levels, each series' fused last level right after
// its last merge.  A dependency always precedes its dependants.
void FFABatch::build_tree(std::vector<uint8_t>& blob, std::vector<SizeType>& offs) {
    float* fold  = m_fold.as<float>();
    auto* coords = m_coords.as<FFACoordFreq>();
    auto buf     = [&](const Series& sr, SizeType level) { return fold + sr.buf_off[level % 2]; };
    const auto ns = m_series.size();
    std::vector<kern::TreeItem> items;
    std::vector<kern::ResolveTask> rt;
    std::vector<kern::FoldTask> ft;
    std::vector<kern::MergeTask> mt;
    std::vector<kern::FinalTask> fl;
    uint32_t ticket = 0;
    SizeType lds    = 0;
    auto add = [&](kern::TreeKind kind, uint32_t nbins, uint32_t task, uint64_t ntiles,
                   uint32_t d0, uint32_t d1, uint32_t lanes) {
        LOKI_REQUIRE(static_cast<uint64_t>(ticket) + ntiles < (1ULL << 32),
                     "FFABatch: too many tiles for one persistent launch");
        kern::TreeItem it{};
        it.kind    = static_cast<uint32_t>(kind);
        it.nbins   = nbins;
        it.task    = task;
        it.ticket0 = ticket;
        it.ntiles  = static_cast<uint32_t>(ntiles);
        it.dep[0]  = d0;
        it.dep[1]  = d1;
        it.lanes   = lanes;
        for (uint32_t d : {d0, d1}) {
            LOKI_REQUIRE(d == kern::kTreeNoDep ||
                             (d < items.size() && items[d].ticket0 + items[d].ntiles <= ticket),
                         "FFABatch: a dependency must precede its dependants");
        }
        ticket += it.ntiles;
        items.push_back(it);
        return static_cast<uint32_t>(items.size() - 1);
    };
    // 1. shift tables of levels 1..L-2 (the last level resolves in-kernel)
    std::vector<std::vector<uint32_t>> res_item(ns), merge_item(ns);
    for (SizeType i = 0; i < ns; ++i) {
        const auto& sr = m_series[i];
        res_item[i].assign(sr.nlevels, kern::kTreeNoDep);
        merge_item[i].assign(sr.nlevels, kern::kTreeNoDep);
        for (SizeType l = 1; l + 1 < sr.nlevels; ++l) {
            const SizeType n = sr.plan.ncoords()[l];
            rt.push_back({resolve_level(sr.plan, l),
                          coords + sr.coord_off + sr.plan.ncoords_offsets64()[l], n, 0});
            res_item[i][l] = add(kern::TreeKind::kResolve, 0, static_cast<uint32_t>(rt.size() - 1),
                                 ceil_div(n, kern::kTreeBlock), kern::kTreeNoDep,
                                 kern::kTreeNoDep, 0);
        }
    }
    // 2. brute folds, the most sample x frequency work first
    std::vector<SizeType> order(ns);
    for (SizeType i = 0; i < ns; ++i) {
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](SizeType a, SizeType b) {
        const auto work = [&](SizeType k) {
            return static_cast<double>(m_series[k].plan.config().get_nsamps()) *
                   static_cast<double>(m_series[k].lists.nfreqs);
        };
        return work(a) > work(b);
    });
    std::vector<uint32_t> fold_item(ns, kern::kTreeNoDep);
    for (const SizeType i : order) {
        const auto& sr = m_series[i];
        const auto& L  = sr.lists;
        kern::FoldTask t{};
        t.ts_e   = ts_e(i);
        t.ts_v   = ts_v(i);
        t.list   = L.list.as<uint32_t>();
        t.base   = L.base.as<uint64_t>();
        t.counts = L.counts.as<uint32_t>();
        t.freqs  = L.freqs.as<double>();
        t.tsamp  = L.tsamp;
        t.t_ref  = L.t_ref;
        t.fold   = buf(sr, 0);
        t.scale  = 1.0F;
        t.seglen = L.seglen;
        t.nfreqs = L.nfreqs;
        t.nbins  = L.nbins;
        t.bps    = kern::brute_fold_blocks_per_segment(L);
        ft.push_back(t);
        const auto kind = L.use_runs ? kern::TreeKind::kFoldUnits : kern::TreeKind::kFoldList;
        lds = std::max(lds, kern::tree_lds_bytes(kind, L.nbins, 0, 0));
        fold_item[i] = add(kind, L.nbins, static_cast<uint32_t>(ft.size() - 1),
                           static_cast<uint64_t>(sr.plan.nsegments()[0]) * t.bps,
                           kern::kTreeNoDep, kern::kTreeNoDep, 0);
    }
    // 3. merge levels and fused last levels, stage by stage
    SizeType max_levels = 0;
    for (const auto& sr : m_series) {
        max_levels = std::max(max_levels, sr.nlevels);
    }
    for (SizeType k = 1; k < max_levels; ++k) {
        for (SizeType i = 0; i < ns; ++i) {
            const auto& sr   = m_series[i];
            const auto& pl   = sr.plan;
            const SizeType L = sr.nlevels;
            const auto nb    = static_cast<uint32_t>(pl.config().get_nbins());
            if (k + 1 < L) {  // merge level k
                kern::MergeTask t{};
                t.in          = buf(sr, k - 1);
                t.out         = buf(sr, k);
                t.coords      = nullptr;
                t.coords_freq = coords + sr.coord_off + pl.ncoords_offsets64()[k];
                t.ncur        = pl.ncoords()[k];
                t.nprev       = pl.ncoords()[k - 1];
                t.ts = kern::make_schedule(t.ncur, pl.param_counts()[k].back(), pl.nsegments()[k],
                                           kern::merge_tile(nb), m_sched.merge_group,
                                           m_sched.merge_run);
                t.scale = 1.0F;
                mt.push_back(t);
                const uint32_t prev = k == 1 ? fold_item[i] : merge_item[i][k - 1];
                merge_item[i][k] = add(kern::TreeKind::kMerge, nb,
                                       static_cast<uint32_t>(mt.size() - 1), t.ts.nblocks, prev,
                                       res_item[i][k], 0);
            } else if (k + 1 == L) {  // fused last level
                const uint32_t lanes = m_sched.score_lanes[__builtin_ctz(nb) - 3];
                kern::FinalTask t{};
                t.in     = buf(sr, L - 2);
                t.scores = m_scores.as<float>() + sr.score_off;
                t.nprev  = pl.ncoords()[L - 2];
                t.lv     = resolve_level(pl, L - 1);
                for (int q = 0; q < 3; ++q) {
                    t.fg.dv[q] = (t.lv.hi[q] - t.lv.lo[q]) / static_cast<double>(t.lv.n_cur[q]);
                }
                t.wt        = kern::make_width_table(sr.widths.data(), sr.widths.size(), nb);
                t.fg.stride = kern::final_lds_stride(nb, t.wt.wmax);
                t.ts = kern::make_schedule(pl.ncoords().back(), t.lv.n_cur[2], 1,
                                           kern::final_wide_rows(lanes), m_sched.final_group,
                                           m_sched.final_run);
                t.unscale = 1.0F;
                fl.push_back(t);
                lds = std::max(lds, kern::tree_lds_bytes(kern::TreeKind::kFinal, nb, lanes,
                                                         t.fg.stride));
                const uint32_t prev = L == 2 ? fold_item[i] : merge_item[i][L - 2];
                add(kern::TreeKind::kFinal, nb, static_cast<uint32_t>(fl.size() - 1),
                    t.ts.nblocks, prev, kern::kTreeNoDep, lanes);
            }
        }
    }
    LOKI_REQUIRE(lds <= 64 * 1024,
                 "FFABatch: a persistent tile needs more than 64 KiB of LDS (more score_lanes "
                 "for the widest rows)");
    offs = {append(blob, items), append(blob, rt), append(blob, ft), append(blob, mt),
            append(blob, fl)};
    m_tree          = kern::TreeArgs{};
    m_tree.nitems   = static_cast<uint32_t>(items.size());
    m_tree.ntickets = ticket;
    m_tree_lds      = lds;
}

void FFABatch::enqueue(hipStream_t s) const {
    if (m_sched.persistent != 0) {
        kern::tree_launch(m_tree, m_tree_blocks, m_tree_lds, s);
        hip_check(hipGetLastError(), "batch persistent launch");
        return;
    }
    const auto* base = static_cast<const uint8_t*>(m_tasks.get());
    for (const auto& ln : m_launches) {
        const void* t = base + ln.task_off;
        switch (ln.stage) {
        case Stage::kResolve:
            kern::resolve_grouped(static_cast<const kern::ResolveTask*>(t), ln.ntasks, ln.nblocks,
                                  1, s);
            break;
        case Stage::kFold:
            kern::brute_fold_grouped(static_cast<const kern::FoldTask*>(t), ln.ntasks,
                                     ln.nblocks, ln.nbins != 0, ln.stride, FoldStorage::kF32, s);
            break;
        case Stage::kMerge:
            kern::merge_grouped(static_cast<const kern::MergeTask*>(t), ln.ntasks, ln.nblocks,
                                ln.nbins, false, FoldStorage::kF32, s);
            break;
        case Stage::kFinal:
            if (ln.lanes == 0) {
                kern::final_grouped(static_cast<const kern::FinalTask*>(t), ln.ntasks,
                                    ln.nblocks, ln.nbins, FoldStorage::kF32, false, true,
                                    ln.stride, s);
            } else {
                kern::final_grouped_wide(static_cast<const kern::FinalTask*>(t), ln.ntasks,
                                         ln.nblocks, ln.nbins, ln.lanes, ln.lds, s);
            }
            break;
        }
        hip_check(hipGetLastError(), "batch launch");
    }
}

void FFABatch::execute(hipStream_t s) {
    DeviceGuard g(m_device);
    if (!m_use_graph) {
        enqueue(s);
        return;
    }
    if (m_exec == nullptr) {
        // capture on a stream of our own (the legacy default stream cannot be captured)
        hipStream_t cap = nullptr;
        hip_check(hipStreamCreate(&cap), "hipStreamCreate");
        hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal),
                  "hipStreamBeginCapture");
        try {
            enqueue(cap);
        } catch (...) {
            hipGraph_t tmp = nullptr;
            (void)hipStreamEndCapture(cap, &tmp);
            if (tmp != nullptr) {
                (void)hipGraphDestroy(tmp);
            }
            (void)hipStreamDestroy(cap);
            throw;
        }
        hip_check(hipStreamEndCapture(cap, &m_graph), "hipStreamEndCapture");
        (void)hipStreamDestroy(cap);
        hip_check(hipGraphInstantiate(&m_exec, m_graph, nullptr, nullptr, 0),
                  "hipGraphInstantiate");
    }
    hip_check(hipGraphLaunch(m_exec, s), "hipGraphLaunch");
}

BatchSchedule FFABatch::autotune(hipStream_t s, int repeats) {
    DeviceGuard g(m_device);
    LOKI_REQUIRE(repeats >= 1, "autotune: repeats must be >= 1");
    const bool graph = m_use_graph;
    m_use_graph      = false;
    hipEvent_t a = nullptr, b = nullptr;
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventCreate(&b), "hipEventCreate");
    // ms per execute() of a candidate, or < 0 if it cannot be built
    auto time_of = [&](const BatchSchedule& c) -> float {
        try {
            set_schedule(c);
        } catch (const std::exception&) {
            return -1.0F;
        }
        enqueue(s);  // warm-up
        hip_check(hipEventRecord(a, s), "hipEventRecord");
        for (int r = 0; r < repeats; ++r) {
            enqueue(s);
        }
        hip_check(hipEventRecord(b, s), "hipEventRecord");
        hip_check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0.0F;
        hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        return ms;
    };
    BatchSchedule best{};
    try {
        // 1. tile orders: merges (last levels at 16 x 16), then last levels
        //    (merges at the winner)
        const BatchSchedule base{};
        float best_ms     = -1.0F;
        BatchSchedule cur = base;
        for (uint32_t mg : {4U, 16U, 64U}) {
            for (uint32_t mr : {4U, 16U, 64U}) {
                BatchSchedule c = base;
                c.merge_group   = mg;
                c.merge_run     = mr;
                const float ms  = time_of(c);
                if (ms >= 0.0F && (best_ms < 0.0F || ms < best_ms)) {
                    best_ms = ms;
                    cur     = c;
                }
            }
        }
        for (uint32_t fg : {4U, 64U}) {
            for (uint32_t fr : {4U, 64U}) {
                BatchSchedule c = cur;
                c.final_group   = fg;
                c.final_run     = fr;
                const float ms  = time_of(c);
                if (ms >= 0.0F && ms < best_ms) {
                    best_ms = ms;
                    cur     = c;
                }
            }
        }
        // 2. boxcar-width schedule of the last levels: lanes per trial, one
        //    nbins class at a time (0: the 64-lane workgroups)
        std::map<uint32_t, int> classes;
        for (const auto& sr : m_series) {
            classes[static_cast<uint32_t>(sr.plan.config().get_nbins())] = 1;
        }
        for (const auto& [nb, unused] : classes) {
            (void)unused;
            const int k = __builtin_ctz(nb) - 3;
            for (uint32_t lanes : {1U, 2U, 4U}) {
                if (!kern::final_wide_supported(nb, lanes)) {
                    continue;
                }
                BatchSchedule c  = cur;
                c.score_lanes[k] = lanes;
                const float ms   = time_of(c);
                if (ms >= 0.0F && ms < best_ms) {
                    best_ms = ms;
                    cur     = c;
                }
            }
        }
        best = cur;
        set_schedule(best);
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        m_use_graph = graph;
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    m_use_graph = graph;
    return best;
}

}  // namespace loki_amd
