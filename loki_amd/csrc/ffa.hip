// FFA level driver on one GPU (restates ffa_cuda.cu:171-196, 409-522 and
// workspace_cuda.cu:83-158 without their per-level host synchronisations):
// shift tables for every level, the brute fold and all merge levels are
// enqueued back to back on the caller's stream.
#include "ffa.hpp"

#include <algorithm>
#include <cmath>
#include <string>

#include "kernels.hpp"
#include "resolve_device.hpp"

namespace loki_amd {

void hip_check(hipError_t err, const char* what) {
    if (err != hipSuccess) {
        throw HipError(std::string(what) + ": " + hipGetErrorString(err));
    }
}

DeviceBuffer::DeviceBuffer(SizeType bytes) { reserve(bytes); }

DeviceBuffer::~DeviceBuffer() {
    if (m_ptr != nullptr) {
        (void)hipFree(m_ptr);
    }
}

DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        if (m_ptr != nullptr) {
            (void)hipFree(m_ptr);
        }
        m_ptr     = o.m_ptr;
        m_bytes   = o.m_bytes;
        o.m_ptr   = nullptr;
        o.m_bytes = 0;
    }
    return *this;
}

void DeviceBuffer::reserve(SizeType bytes) {
    if (bytes <= m_bytes && m_ptr != nullptr) {
        return;
    }
    if (m_ptr != nullptr) {
        hip_check(hipFree(m_ptr), "hipFree");
        m_ptr   = nullptr;
        m_bytes = 0;
    }
    if (bytes == 0) {
        return;
    }
    hip_check(hipMalloc(&m_ptr, bytes), "hipMalloc");
    m_bytes = bytes;
}

FFAWorkspace::~FFAWorkspace() {
    if (m_done != nullptr) {
        (void)hipEventDestroy(m_done);
    }
}

void FFAWorkspace::begin_use(hipStream_t s) {
    if (m_recorded && m_last != s) {
        hip_check(hipStreamWaitEvent(s, m_done, 0), "workspace: wait for the previous stream");
    }
}

void FFAWorkspace::end_use(hipStream_t s) {
    if (m_done == nullptr) {
        hip_check(hipEventCreateWithFlags(&m_done, hipEventDisableTiming), "hipEventCreate");
    }
    hip_check(hipEventRecord(m_done, s), "workspace: record use");
    m_last     = s;
    m_recorded = true;
}

DeviceGuard::DeviceGuard(int device) {
    int count = 0;
    hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    LOKI_REQUIRE(device >= 0 && device < count, "invalid device_id");
    hip_check(hipGetDevice(&m_prev), "hipGetDevice");
    if (m_prev != device) {
        hip_check(hipSetDevice(device), "hipSetDevice");
        m_changed = true;
    }
}

DeviceGuard::~DeviceGuard() {
    if (m_changed) {
        (void)hipSetDevice(m_prev);
    }
}

LaunchTimer::~LaunchTimer() {
    for (auto& sl : m_slots) {
        (void)hipEventDestroy(sl.a);
        (void)hipEventDestroy(sl.b);
    }
}

LaunchTimer::LaunchTimer(LaunchTimer&& o) noexcept
    : m_enabled(o.m_enabled), m_slots(std::move(o.m_slots)), m_used(o.m_used) {
    o.m_slots.clear();
    o.m_used = 0;
}

int LaunchTimer::start(KernelKind kind, int level, hipStream_t s) {
    if (!m_enabled) {
        return -1;
    }
    if (m_used == m_slots.size()) {
        Slot sl;
        hip_check(hipEventCreate(&sl.a), "hipEventCreate");
        hip_check(hipEventCreate(&sl.b), "hipEventCreate");
        m_slots.push_back(sl);
    }
    Slot& sl = m_slots[m_used];
    sl.kind  = kind;
    sl.level = level;
    hip_check(hipEventRecord(sl.a, s), "hipEventRecord");
    return static_cast<int>(m_used++);
}

void LaunchTimer::stop(int slot, hipStream_t s) {
    if (slot < 0) {
        return;
    }
    hip_check(hipEventRecord(m_slots[static_cast<size_t>(slot)].b, s), "hipEventRecord");
}

std::vector<KernelTiming> LaunchTimer::collect() {
    std::vector<KernelTiming> out;
    for (size_t i = 0; i < m_used; ++i) {
        hip_check(hipEventSynchronize(m_slots[i].b), "hipEventSynchronize");
        float ms = 0.0F;
        hip_check(hipEventElapsedTime(&ms, m_slots[i].a, m_slots[i].b), "hipEventElapsedTime");
        out.push_back({m_slots[i].kind, m_slots[i].level, ms});
    }
    return out;
}

kern::ResolveLevel resolve_level(const FFAPlan& plan, SizeType l) {
    const auto& cfg  = plan.config();
    const auto lims  = cfg.get_param_limits();
    const SizeType n = plan.n_params();
    kern::ResolveLevel lv{};
    for (int k = 0; k < 3; ++k) {
        lv.n_cur[k]  = 1;
        lv.n_prev[k] = 1;
        lv.lo[k]     = 0.0;
        lv.hi[k]     = 1.0;
    }
    const SizeType use = n < 3 ? n : 3;
    const SizeType po  = n - use;
    for (SizeType p = 0; p < po; ++p) {
        LOKI_REQUIRE(plan.param_counts()[l][p] == 1 && plan.param_counts()[l - 1][p] == 1,
                     "crackle/snap grids must have a single point for the FFA resolver");
    }
    for (SizeType k = 0; k < use; ++k) {
        const SizeType p   = po + k;
        const SizeType dst = 3 - use + k;
        LOKI_REQUIRE(plan.param_counts()[l][p] < (1ULL << 32), "grid count exceeds 2^32");
        lv.n_cur[dst]  = static_cast<uint32_t>(plan.param_counts()[l][p]);
        lv.n_prev[dst] = static_cast<uint32_t>(plan.param_counts()[l - 1][p]);
        lv.lo[dst]     = lims[p].min;
        lv.hi[dst]     = lims[p].max;
    }
    lv.tseg_prev = std::ldexp(cfg.get_tseg_brute(), static_cast<int>(l) - 1);
    lv.nbins     = static_cast<uint32_t>(cfg.get_nbins());
    lv.np        = static_cast<uint32_t>(use);
    return lv;
}

namespace {

void check_launch(const char* what) { hip_check(hipGetLastError(), what); }

/// Resolver terms of grid row `row` (np 2 or 3), one side; dv as in
/// FinalGrid.
kern::RowTerms row_terms_of(const kern::ResolveLevel& lv, const double* dv, uint32_t row,
                            int side) {
    return lv.np == 2 ? kern::row_terms<2>(lv, dv, row, side)
                      : kern::row_terms<3>(lv, dv, row, side);
}

/// Every grid row's (tail, head) terms of one level, np 2 or 3.
void level_row_terms(const kern::ResolveLevel& lv, std::vector<kern::RowTerms>& out) {
    double dv[3];
    for (int k = 0; k < 3; ++k) {
        dv[k] = (lv.hi[k] - lv.lo[k]) / static_cast<double>(lv.n_cur[k]);
    }
    const uint64_t nrows = static_cast<uint64_t>(lv.n_cur[0]) * lv.n_cur[1];
    for (uint64_t r = 0; r < nrows; ++r) {
        for (int side = 0; side < 2; ++side) {
            out.push_back(row_terms_of(lv, dv, static_cast<uint32_t>(r), side));
        }
    }
}

}  // namespace

FFAExecutor::FFAExecutor(const SearchConfig& cfg, FoldKind kind, int device_id,
                         std::shared_ptr<FFAWorkspace> ws)
    : m_plan(cfg, kind),
      m_device(device_id),
      m_freq_only(cfg.get_nparams() == 1),
      m_fourier(kind == FoldKind::kFourier),
      m_lossy(kind == FoldKind::kFourier && cfg.get_nbins() > cfg.get_nbins_min_lossy_bf()),
      m_ws(ws ? std::move(ws) : std::make_shared<FFAWorkspace>()) {
    DeviceGuard g(device_id);
    LOKI_REQUIRE(cfg.get_nbins() <= 1024, "nbins must be <= 1024");
    LOKI_REQUIRE(!m_fourier || m_lossy || (cfg.get_nbins() >= 2 && cfg.get_nbins() <= 256),
                 "Fourier brute fold needs 2 <= nbins <= 256 (larger nbins use the lossy "
                 "time-domain init: nbins > nbins_min_lossy_bf)");
    for (SizeType l = 1; l < m_plan.n_levels(); ++l) {
        LOKI_REQUIRE(m_plan.ncoords()[l - 1] < (1ULL << 32),
                     "parent grid exceeds 2^32 coordinates (u32 indices)");
    }
    // level-0 scale of reduced-precision storage: bucket sums of about
    // seglen/nbins samples are stored near unit magnitude
    for (SizeType per = m_plan.segment_lens()[0] / cfg.get_nbins(); per > 1; per >>= 1) {
        ++m_k0;
    }
}

SizeType FFAExecutor::coord_record_bytes() const noexcept {
    return m_freq_only ? sizeof(FFACoordFreq) : sizeof(FFACoord);
}

void FFAExecutor::resolve_coordinates(void* coords_dev, hipStream_t s, SizeType last_level) {
    auto* base = static_cast<char*>(coords_dev);
    const SizeType top = std::min(last_level, m_plan.n_levels() - 1);
    if (!m_freq_only && !m_rt_ready) {
        // the frequency-independent resolver terms of every level's grid rows
        m_rt_host.clear();
        m_rt_off.assign(m_plan.n_levels(), 0);
        for (SizeType l = 1; l < m_plan.n_levels(); ++l) {
            m_rt_off[l] = m_rt_host.size();
            level_row_terms(resolve_level(m_plan, l), m_rt_host);
        }
        m_rt.reserve(std::max<SizeType>(m_rt_host.size(), 1) * sizeof(kern::RowTerms));
        hip_check(hipMemcpyAsync(m_rt.get(), m_rt_host.data(),
                                 m_rt_host.size() * sizeof(kern::RowTerms),
                                 hipMemcpyHostToDevice, s),
                  "copy resolver row terms");
        m_rt_ready = true;
    }
    for (SizeType l = 1; l <= top; ++l) {
        const auto lv      = resolve_level(m_plan, l);
        const SizeType off = m_plan.ncoords_offsets64()[l];
        const SizeType n   = m_plan.ncoords()[l];
        const int slot     = m_timer.start(KernelKind::kResolve, static_cast<int>(l), s);
        if (m_freq_only) {
            kern::resolve_freq(lv, reinterpret_cast<FFACoordFreq*>(base) + off, n, s);
        } else {
            kern::resolve_poly(lv, m_rt.as<kern::RowTerms>() + m_rt_off[l],
                               reinterpret_cast<FFACoord*>(base) + off, n, s);
        }
        check_launch("resolver launch");
        m_timer.stop(slot, s);
    }
}

void FFAExecutor::prepare_brute_fold(hipStream_t s) {
    if (m_buckets_ready) {
        return;
    }
    const auto freqs      = m_plan.compute_param_grid(0).back();
    const SizeType nfreqs = freqs.size();
    const SizeType seglen = m_plan.segment_lens()[0];
    const SizeType nbins  = m_plan.config().get_nbins();
    const double t_ref    = m_freq_only ? 0.0 : m_plan.tsegments()[0] / 2.0;
    LOKI_REQUIRE(seglen < (1ULL << 32) && nfreqs < (1ULL << 32), "brute fold too large");
    m_bf_freqs.reserve(nfreqs * sizeof(double));
    hip_check(hipMemcpyAsync(m_bf_freqs.get(), freqs.data(), nfreqs * sizeof(double),
                             hipMemcpyHostToDevice, s),
              "copy brute-fold frequencies");
    if (!m_fourier || m_lossy) {
        // synchronises s (so `freqs` outlives the copy)
        m_bf = kern::build_brute_fold_lists(
            m_bf_freqs.as<double>(), static_cast<uint32_t>(nfreqs), static_cast<uint32_t>(seglen),
            static_cast<uint32_t>(nbins), m_plan.config().get_tsamp(), t_ref, s);
    } else {
        hip_check(hipStreamSynchronize(s), "copy brute-fold frequencies");
    }
    m_buckets_ready = true;
}

void FFAExecutor::set_fold_storage(FoldStorage st) {
    LOKI_REQUIRE(!m_fourier || st == FoldStorage::kF32,
                 "reduced-precision fold storage is available for the time-domain FFA only");
    if (st == FoldStorage::kBFP16) {
        const SizeType nb = m_plan.config().get_nbins();
        LOKI_REQUIRE(nb >= 4 && nb <= 256 && (nb & (nb - 1)) == 0,
                     "block floating-point storage needs a power-of-two nbins in [4, 256]");
    }
    m_storage = st;
}

FoldStorage FFAExecutor::stored_as(SizeType l, SizeType last_level, bool last_fp32) const {
    if (l == last_level && last_fp32) {
        return FoldStorage::kF32;
    }
    // block floating point keeps the brute fold (level 0) in FP32
    if (l == 0 && m_storage == FoldStorage::kBFP16) {
        return FoldStorage::kF32;
    }
    return m_storage;
}

SizeType FFAExecutor::stored_bytes(SizeType last_level, bool last_fp32) const {
    SizeType m = 0;
    const SizeType nbins = m_plan.config().get_nbins();
    for (SizeType l = 0; l <= last_level; ++l) {
        m = std::max(m, static_cast<SizeType>(level_bytes(stored_as(l, last_level, last_fp32),
                                                          m_plan.level_size(l), nbins)));
    }
    return m;
}

void FFAExecutor::init_level(const float* ts_e, const float* ts_v, void* out, void* scratch,
                             bool fp32_out, hipStream_t s) {
    const SizeType nbins = m_plan.config().get_nbins();
    const SizeType nseg  = m_plan.nsegments()[0];
    const int bslot      = m_timer.start(KernelKind::kBruteFold, 0, s);
    if (!m_fourier) {
        // FP16/BF16 level 0 is stored at 2^-k0; FP32 (also under block
        // floating point) at its true scale (the kernel ignores the scale)
        const FoldStorage so = stored_as(0, 0, fp32_out);
        kern::brute_fold_time(ts_e, ts_v, nseg, m_bf, out, so, std::ldexp(1.0F, -m_k0), s);
        check_launch("brute fold launch");
        m_timer.stop(bslot, s);
        return;
    }
    if (!m_lossy) {
        const double t_ref = m_freq_only ? 0.0 : m_plan.tsegments()[0] / 2.0;
        kern::brute_fold_fourier(ts_e, ts_v, nseg, static_cast<uint32_t>(m_plan.segment_lens()[0]),
                                 m_bf_freqs.as<double>(),
                                 static_cast<uint32_t>(m_plan.param_counts()[0].back()),
                                 static_cast<uint32_t>(nbins), m_plan.config().get_tsamp(),
                                 t_ref, static_cast<float2*>(out), s, m_dft_split);
        check_launch("Fourier brute fold launch");
        m_timer.stop(bslot, s);
        return;
    }
    // Lossy init (ffa_cuda.cu:370-398): time brute fold into the partner
    // buffer, then a forward real DFT of every (segment, freq, component) row.
    auto* tf = static_cast<float*>(scratch);
    kern::brute_fold_time(ts_e, ts_v, nseg, m_bf, tf, FoldStorage::kF32, 1.0F, s);
    check_launch("brute fold launch");
    m_timer.stop(bslot, s);
    const SizeType nrows = nseg * m_bf.nfreqs * 2;
    const int tslot      = m_timer.start(KernelKind::kTransform, 0, s);
    kern::rfft_rows(tf, static_cast<float2*>(out), nrows, static_cast<uint32_t>(nbins), s);
    check_launch("rfft launch");
    m_timer.stop(tslot, s);
}

void* FFAExecutor::run_levels(const float* ts_e, const float* ts_v, void* buf_even,
                              void* buf_odd, SizeType last_level, bool last_fp32,
                              hipStream_t s, SizeType coords_level) {
    // Level k lands in buf_even for even k and in buf_odd for odd k (the
    // reference picks its ping-pong start side the same way so the last level
    // lands in the caller's buffer, ffa.cpp:261-276).
    LOKI_REQUIRE(last_level < m_plan.n_levels(), "level out of range");
    m_timer.reset();
    m_ws->coords.reserve(m_plan.coord_size() * coord_record_bytes());
    prepare_brute_fold(s);
    resolve_coordinates(m_ws->coords.get(), s, std::max(last_level, coords_level));

    void* even           = buf_even;
    void* odd            = buf_odd;
    const SizeType nbins = m_plan.config().get_nbins();
    init_level(ts_e, ts_v, even, odd, last_fp32 && last_level == 0, s);

    auto tables = [&](SizeType l, const FFACoord*& cp, const FFACoordFreq*& cfp) {
        const SizeType off = m_plan.ncoords_offsets64()[l];
        cp                 = nullptr;
        cfp                = nullptr;
        if (m_freq_only) {
            cfp = m_ws->coords.as<FFACoordFreq>() + off;
        } else {
            cp = m_ws->coords.as<FFACoord>() + off;
        }
    };
    // Pass plan: levels l and l + 1 merged in one two-level pass where allowed
    // (time domain, both stored in the tree's storage), pairing from the top
    // down; the number of two-level passes is kept even, so that level k
    // still lands in buf_even for even k and in buf_odd for odd k.
    std::vector<uint8_t> two(last_level + 2, 0);  // two[l]: levels l, l+1 in one pass
    if (!m_fourier && m_two_level && kern::merge2_supported(static_cast<uint32_t>(nbins))) {
        std::vector<SizeType> starts;
        for (SizeType top = last_level; top >= 2;) {
            const SizeType l     = top - 1;
            const FoldStorage si = stored_as(l - 1, last_level, last_fp32);
            const FoldStorage sb = stored_as(l, last_level, last_fp32);
            const FoldStorage so = stored_as(top, last_level, last_fp32);
            const bool to32      = last_fp32 && top == last_level;
            // only where the tree does not grow: a pass whose output level is
            // larger than its input level (the accel/jerk grid refining) gathers
            // its four input rows from a level half its size and measured slower
            // than two single passes (2^23-sample accel search, levels 10-11:
            // 1.97 ms against 0.47 + 1.28 ms)
            const bool grows = m_plan.level_size(top) > m_plan.level_size(l - 1) + m_plan.level_size(l - 1) / 4;
            if (to32 || grows || sb != so ||
                !(si == so || (si == FoldStorage::kF32 && so == FoldStorage::kBFP16))) {
                top -= 1;  // this pair cannot be merged in one pass; try one lower
                continue;
            }
            starts.push_back(l);
            top -= 2;
        }
        if (starts.size() % 2 == 1) {
            starts.pop_back();  // the lowest pair
        }
        for (const SizeType l : starts) {
            two[l] = 1;
        }
    }
    void* cur   = even;
    void* other = odd;
    for (SizeType l = 1; l <= last_level;) {
        const FFACoord* cp      = nullptr;
        const FFACoordFreq* cfp = nullptr;
        tables(l, cp, cfp);
        if (two[l] != 0) {
            const FFACoord* cp2      = nullptr;
            const FFACoordFreq* cfp2 = nullptr;
            tables(l + 1, cp2, cfp2);
            const int slot = m_timer.start(KernelKind::kMerge, static_cast<int>(l + 1), s);
            kern::merge_time2(cur, other, cp, cfp, cp2, cfp2, m_plan.ncoords()[l + 1],
                              m_plan.ncoords()[l], m_plan.ncoords()[l - 1],
                              m_plan.nsegments()[l + 1], m_plan.param_counts()[l + 1].back(),
                              static_cast<uint32_t>(nbins), stored_as(l - 1, last_level, last_fp32),
                              stored_as(l + 1, last_level, last_fp32), 0.5F, s);
            check_launch("two-level merge launch");
            m_timer.stop(slot, s);
            l += 2;
        } else {
            const int slot = m_timer.start(KernelKind::kMerge, static_cast<int>(l), s);
            if (m_fourier) {
                kern::merge_fourier(static_cast<const float2*>(cur), static_cast<float2*>(other),
                                    cp, cfp, m_plan.ncoords()[l], m_plan.ncoords()[l - 1],
                                    m_plan.nsegments()[l], static_cast<uint32_t>(nbins), s);
            } else {
                // FP16/BF16: (tail + head) * 0.5 keeps level l at 2^-(k0 + l); a last
                // level returned in FP32 is unscaled by 2^(k0 + l - 1) instead.
                // FP32 and block floating point are unscaled (scale unused).
                const bool to32      = last_fp32 && l == last_level;
                const FoldStorage si = stored_as(l - 1, last_level, last_fp32);
                const FoldStorage so = stored_as(l, last_level, last_fp32);
                const float scale =
                    to32 ? std::ldexp(1.0F, m_k0 + static_cast<int>(l) - 1) : 0.5F;
                kern::merge_time(cur, other, cp, cfp, m_plan.ncoords()[l],
                                 m_plan.ncoords()[l - 1], m_plan.nsegments()[l],
                                 m_plan.param_counts()[l].back(), static_cast<uint32_t>(nbins),
                                 si, so, scale, s);
            }
            check_launch("merge launch");
            m_timer.stop(slot, s);
            l += 1;
        }
        std::swap(cur, other);
    }
    LOKI_REQUIRE(cur == ((last_level % 2 == 0) ? even : odd), "merge passes: wrong landing buffer");
    return cur;
}

SizeType FFAExecutor::max_level_size(SizeType last_level) const {
    SizeType m = 0;
    for (SizeType l = 0; l <= last_level; ++l) {
        m = std::max(m, m_plan.level_size(l));
    }
    return m;
}

void FFAExecutor::execute(const float* ts_e, const float* ts_v, void* fold, SizeType fold_len,
                          hipStream_t s) {
    DeviceGuard g(m_device);
    LOKI_REQUIRE(ts_e != nullptr && ts_v != nullptr && fold != nullptr,
                 "FFA::execute: null buffer");
    LOKI_REQUIRE(fold_len == m_plan.buffer_size(),
                 "FFA::execute: fold must have size buffer_size");
    const SizeType L = m_plan.n_levels();
    m_ws->begin_use(s);
    m_ws->internal.reserve(m_plan.buffer_size() * elem_bytes());
    if ((L - 1) % 2 == 0) {
        run_levels(ts_e, ts_v, fold, m_ws->internal.get(), L - 1, true, s);
    } else {
        run_levels(ts_e, ts_v, m_ws->internal.get(), fold, L - 1, true, s);
    }
    m_ws->end_use(s);
}

void FFAExecutor::execute_time(const float* ts_e, const float* ts_v, float* fold,
                               SizeType fold_len, hipStream_t s) {
    DeviceGuard g(m_device);
    LOKI_REQUIRE(m_fourier, "FFA::execute: the time-domain output overload needs a Fourier plan");
    LOKI_REQUIRE(ts_e != nullptr && ts_v != nullptr && fold != nullptr,
                 "FFA::execute: null buffer");
    LOKI_REQUIRE(fold_len == 2 * m_plan.buffer_size(),
                 "FFA::execute: fold must have size 2*buffer_size_fourier");
    const SizeType L = m_plan.n_levels();
    // The caller's buffer (buffer_size complex elements) is the ping-pong
    // partner; the last level lands in the internal buffer and is returned to
    // the time domain into the caller's buffer (ffa_cuda.cu:248-290).
    m_ws->begin_use(s);
    m_ws->internal.reserve(m_plan.buffer_size() * elem_bytes());
    const void* fin = ((L - 1) % 2 == 0)
                          ? run_levels(ts_e, ts_v, m_ws->internal.get(), fold, L - 1, true, s)
                          : run_levels(ts_e, ts_v, fold, m_ws->internal.get(), L - 1, true, s);
    const SizeType nbins = m_plan.config().get_nbins();
    const int slot       = m_timer.start(KernelKind::kTransform, static_cast<int>(L - 1), s);
    kern::irfft_rows(static_cast<const float2*>(fin), fold, m_plan.fold_size_time() / nbins,
                     static_cast<uint32_t>(nbins), s);
    check_launch("irfft launch");
    m_timer.stop(slot, s);
    m_ws->end_use(s);
}

bool FFAExecutor::fused_scores_supported() const {
    return !m_fourier &&
           kern::final_merge_score_supported(
               static_cast<uint32_t>(m_plan.config().get_nbins())) &&
           m_plan.nsegments().back() == 1;
}

void FFAExecutor::set_fuse_depth(int depth) {
    LOKI_REQUIRE(depth == 1 || depth == 2, "fuse depth must be 1 or 2");
    m_fuse_depth = depth;
}

bool FFAExecutor::final2_usable() const {
    return !m_fourier && kern::final2_supported(static_cast<uint32_t>(m_plan.config().get_nbins())) &&
           m_plan.nsegments().back() == 1 && m_plan.n_levels() >= 2;
}

namespace {

FFACoord resolve_rc(const kern::ResolveLevel& lv, const double* dv, uint32_t row,
                          uint32_t jf) {
    switch (lv.np) {
    case 1: return kern::resolve_pair_rc<1>(lv, dv, row, jf);
    case 2: return kern::resolve_pair_rc<2>(lv, dv, row, jf);
    default: return kern::resolve_pair_rc<3>(lv, dv, row, jf);
    }
}

}  // namespace

void FFAExecutor::prepare_final2(hipStream_t s) {
    if (m_f2_ready) {
        return;
    }
    const SizeType L   = m_plan.n_levels();
    const auto lv      = resolve_level(m_plan, L - 1);
    double dv[3];
    for (int k = 0; k < 3; ++k) {
        dv[k] = (lv.hi[k] - lv.lo[k]) / static_cast<double>(lv.n_cur[k]);
    }
    const uint32_t n_f    = lv.n_cur[2];
    const uint32_t n_pf   = lv.n_prev[2];
    const SizeType n_rows = m_plan.ncoords()[L - 1] / n_f;
    LOKI_REQUIRE(n_rows < (1ULL << 32), "final level: too many grid rows");
    // parent rows (tail, head) of every final row: they depend on the row's
    // (jerk, accel) cell only
    std::vector<uint64_t> pt(n_rows), ph(n_rows);
    for (SizeType r = 0; r < n_rows; ++r) {
        const auto c = resolve_rc(lv, dv, static_cast<uint32_t>(r), 0);
        pt[r]        = c.i_tail / n_pf;
        ph[r]        = c.i_head / n_pf;
    }
    // runs of consecutive rows with the same parents
    std::vector<std::pair<uint32_t, uint32_t>> groups;  // (row0, size)
    for (SizeType r = 0; r < n_rows;) {
        SizeType e = r + 1;
        while (e < n_rows && pt[e] == pt[r] && ph[e] == ph[r]) {
            ++e;
        }
        groups.emplace_back(static_cast<uint32_t>(r), static_cast<uint32_t>(e - r));
        r = e;
    }
    // chunks of up to 16 rows of one group: a block stages the group's parent
    // rows once for all its rows (its four waves take 4 rows each)
    constexpr uint32_t kChunkRows = 16;
    m_f2_rows_per_block = 1;
    m_f2_chunks_host.clear();
    for (const auto& g : groups) {
        for (uint32_t o = 0; o < g.second; o += kChunkRows) {
            const uint32_t n = std::min(kChunkRows, g.second - o);
            m_f2_chunks_host.push_back({g.first + o, n});
            m_f2_rows_per_block = std::max(m_f2_rows_per_block, n);
        }
    }
    const uint32_t n_ft    = (n_f + 63) / 64;
    const uint64_t nblocks = static_cast<uint64_t>(m_f2_chunks_host.size()) * n_ft;
    LOKI_REQUIRE(nblocks < (1ULL << 31), "final level: grid too large");
    m_f2_nblocks = static_cast<uint32_t>(nblocks);
    m_f2_chunks.reserve(m_f2_chunks_host.size() * sizeof(kern::RowChunk));
    hip_check(hipMemcpyAsync(m_f2_chunks.get(), m_f2_chunks_host.data(),
                             m_f2_chunks_host.size() * sizeof(kern::RowChunk),
                             hipMemcpyHostToDevice, s),
              "copy final-level row chunks");
    // the frequency-independent resolver terms of every grid row, formed once
    // here with the kernel's own (host and device) code
    m_f2_rows_host.clear();
    if (lv.np >= 2) {
        m_f2_rows_host.resize(2 * n_rows);
        for (SizeType r = 0; r < n_rows; ++r) {
            for (int side = 0; side < 2; ++side) {
                m_f2_rows_host[2 * r + side] = row_terms_of(lv, dv, static_cast<uint32_t>(r), side);
            }
        }
        m_f2_rows.reserve(m_f2_rows_host.size() * sizeof(kern::RowTerms));
        hip_check(hipMemcpyAsync(m_f2_rows.get(), m_f2_rows_host.data(),
                                 m_f2_rows_host.size() * sizeof(kern::RowTerms),
                                 hipMemcpyHostToDevice, s),
                  "copy final-level row terms");
    }
    // parent rows staged per side: the exact widest range of any block, from
    // the device resolver itself (one pass per plan; synchronises s)
    kern::Final2Args a{};
    a.lv = lv;
    for (int k = 0; k < 3; ++k) {
        a.fg.dv[k] = dv[k];
    }
    a.chunks   = m_f2_chunks.as<kern::RowChunk>();
    a.n_chunks = static_cast<uint32_t>(m_f2_chunks_host.size());
    DeviceBuffer d_cap(sizeof(uint32_t));
    kern::zero_counter(d_cap.as<uint32_t>(), s);
    kern::final2_ranges(a, m_f2_nblocks, d_cap.as<uint32_t>(), s);
    check_launch("final-level range launch");
    uint32_t cap = 0;
    hip_check(hipMemcpyAsync(&cap, d_cap.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s),
              "copy final-level range");
    hip_check(hipStreamSynchronize(s), "final-level range");
    m_f2_cap = cap;
    m_f2_ok  = cap > 0 && kern::final2_lds_bytes(static_cast<uint32_t>(m_plan.config().get_nbins()),
                                                  cap) <= 64U * 1024U;
    m_f2_ready = true;
}

void FFAExecutor::execute_scores(const float* ts_e, const float* ts_v,
                                 std::span<const SizeType> widths, float* scores,
                                 SizeType scores_len, bool find_max, hipStream_t s,
                                 const kern::FilterOut* fo) {
    DeviceGuard g(m_device);
    m_ws->begin_use(s);
    scores_levels(ts_e, ts_v, widths, scores, scores_len, find_max, s, fo);
    m_ws->end_use(s);
}

void FFAExecutor::scores_levels(const float* ts_e, const float* ts_v,
                                std::span<const SizeType> widths, float* scores,
                                SizeType scores_len, bool find_max, hipStream_t s,
                                const kern::FilterOut* fo) {
    LOKI_REQUIRE(m_plan.nsegments().back() == 1,
                 "compute_ffa_scores: nsegments must be 1 for scores");
    LOKI_REQUIRE(ts_e != nullptr && ts_v != nullptr, "FFA::execute: null buffer");
    const SizeType L       = m_plan.n_levels();
    const SizeType nbins   = m_plan.config().get_nbins();
    const SizeType ncoords = m_plan.ncoords().back();
    const auto wt          = kern::make_width_table(widths.data(), widths.size(), nbins);
    if (fo == nullptr) {
        LOKI_REQUIRE(scores != nullptr, "scores buffer is null");
        LOKI_REQUIRE(scores_len == (find_max ? ncoords : ncoords * widths.size()),
                     "scores must have size ncoords * nwidths (or ncoords for the max)");
    } else {
        LOKI_REQUIRE(fo->count != nullptr &&
                         (fo->capacity == 0 || (fo->indices != nullptr && fo->snr != nullptr)),
                     "score filter: null output buffer");
        kern::zero_counter(fo->count, s);
    }
    if (m_fused && final2_usable()) {
        prepare_final2(s);
    }
    if (m_fused && final2_usable() && m_f2_ok) {
        // One lane per trial: the last level fused with its shift tables and
        // the S/N; with depth 2 its parent level is merged on chip as well.
        // Neither the final fold cube nor (depth 2) the parent level is stored.
        const bool grand    = m_fuse_depth == 2 && L >= 3;
        const SizeType top  = grand ? L - 3 : L - 2;  // last level through HBM
        const SizeType need = stored_bytes(top, false);
        m_ws->internal.reserve(need);
        m_ws->result.reserve(need);
        const void* stored =
            run_levels(ts_e, ts_v, m_ws->result.get(), m_ws->internal.get(), top, false, s, L - 2);
        kern::Final2Args a{};
        a.in   = stored;
        if (stored_as(top, top, false) == FoldStorage::kBFP16) {
            a.ein = reinterpret_cast<const uint16_t*>(static_cast<const char*>(stored) +
                                                      2 * m_plan.level_size(top));
        }
        a.n_p1 = m_plan.ncoords()[L - 2];
        a.n_p2 = grand ? m_plan.ncoords()[L - 3] : 0;
        if (grand) {
            const SizeType off = m_plan.ncoords_offsets64()[L - 2];
            if (m_freq_only) {
                a.coords1f = m_ws->coords.as<FFACoordFreq>() + off;
            } else {
                a.coords1 = m_ws->coords.as<FFACoord>() + off;
            }
        }
        a.lv = resolve_level(m_plan, L - 1);
        for (int k = 0; k < 3; ++k) {
            a.fg.dv[k] = (a.lv.hi[k] - a.lv.lo[k]) / static_cast<double>(a.lv.n_cur[k]);
        }
        a.wt             = wt;
        a.scores         = scores;
        a.fo             = fo != nullptr ? *fo : kern::FilterOut{};
        a.chunks         = m_f2_chunks.as<kern::RowChunk>();
        a.n_chunks       = static_cast<uint32_t>(m_f2_chunks_host.size());
        a.rows           = m_f2_rows_host.empty() ? nullptr : m_f2_rows.as<kern::RowTerms>();
        a.rows_per_block = m_f2_rows_per_block;
        a.cap            = m_f2_cap;
        a.grand          = grand ? 1U : 0U;
        a.unscale        = std::ldexp(1.0F, m_k0 + static_cast<int>(L) - 2);
        a.scale1         = 0.5F;
        if (fo != nullptr) {
            LOKI_REQUIRE(ncoords * (find_max ? 1U : widths.size()) <= 0xFFFFFFFFULL,
                         "score filter: trial index exceeds 32 bits");
        }
        const int slot = m_timer.start(KernelKind::kMergeScore, static_cast<int>(L - 1), s);
        // the stored level may be FP32 (the brute fold) while the parent level
        // merged on chip is block floating point
        a.pbfp = grand && stored_as(L - 2, L - 2, false) == FoldStorage::kBFP16 ? 1U : 0U;
        kern::final2_merge_score(a, stored_as(top, top, false), m_f2_nblocks, find_max,
                                 m_exact_norm, s);
        check_launch("final-level merge+score launch");
        m_timer.stop(slot, s);
        return;
    }
    if (m_fused && fused_scores_supported()) {
        // Levels 0..L-2 through HBM, the last level fused with its shift
        // tables and the S/N: the final fold cube is never stored.
        const SizeType need = stored_bytes(L - 2, false);
        m_ws->internal.reserve(need);
        m_ws->result.reserve(need);
        const void* parent =
            run_levels(ts_e, ts_v, m_ws->result.get(), m_ws->internal.get(), L - 2, false, s);
        const auto lv  = resolve_level(m_plan, L - 1);
        const int slot = m_timer.start(KernelKind::kMergeScore, static_cast<int>(L - 1), s);
        kern::final_merge_score(parent, stored_as(L - 2, L - 2, false),
                                std::ldexp(1.0F, m_k0 + static_cast<int>(L) - 2),
                                ncoords, m_plan.ncoords()[L - 2], lv, wt, scores, find_max,
                                m_exact_norm, s, fo);
        check_launch("final-level merge+score launch");
        m_timer.stop(slot, s);
        return;
    }
    m_ws->internal.reserve(m_plan.buffer_size() * elem_bytes());
    m_ws->result.reserve(m_plan.buffer_size() * elem_bytes());
    const void* last =
        run_levels(ts_e, ts_v, m_ws->result.get(), m_ws->internal.get(), L - 1, true, s);
    const float* fin = static_cast<const float*>(last);
    if (m_fourier && m_fused && kern::irfft_score_supported(static_cast<uint32_t>(nbins))) {
        const int slot =
            m_timer.start(KernelKind::kTransformScore, static_cast<int>(L - 1), s);
        kern::irfft_score(static_cast<const float2*>(last), ncoords, static_cast<uint32_t>(nbins),
                          wt, scores, find_max, m_exact_norm, s, fo);
        check_launch("irfft+score launch");
        m_timer.stop(slot, s);
        return;
    }
    if (m_fourier) {
        // Back to the time domain into the other ping-pong buffer: it holds
        // buffer_size complex = 2*buffer_size floats >= fold_size_time.
        void* other     = last == m_ws->result.get() ? m_ws->internal.get() : m_ws->result.get();
        auto* tf        = static_cast<float*>(other);
        const int tslot = m_timer.start(KernelKind::kTransform, static_cast<int>(L - 1), s);
        kern::irfft_rows(static_cast<const float2*>(last), tf, m_plan.fold_size_time() / nbins,
                         static_cast<uint32_t>(nbins), s);
        check_launch("irfft launch");
        m_timer.stop(tslot, s);
        fin = tf;
    }
    const int slot = m_timer.start(KernelKind::kScore, static_cast<int>(L - 1), s);
    kern::snr_boxcar_3d(fin, ncoords, static_cast<uint32_t>(nbins), wt, scores, find_max, s, fo);
    check_launch("score launch");
    m_timer.stop(slot, s);
}

void brute_fold_time_device(const float* ts_e, const float* ts_v, SizeType nsamps,
                            std::span<const double> freqs, SizeType segment_len, SizeType nbins,
                            double tsamp, double t_ref, float* fold, hipStream_t s) {
    LOKI_REQUIRE(!freqs.empty(), "BruteFold: Frequency array is empty");
    LOKI_REQUIRE(segment_len > 0 && nsamps % segment_len == 0,
                 "BruteFold: Number of samples is not a multiple of segment length");
    LOKI_REQUIRE(nbins >= 1 && nbins <= 1024, "BruteFold: nbins must be in [1, 1024]");
    LOKI_REQUIRE(segment_len < (1ULL << 32), "BruteFold: segment too long");
    DeviceBuffer d_freqs(freqs.size() * sizeof(double));
    hip_check(hipMemcpyAsync(d_freqs.get(), freqs.data(), freqs.size() * sizeof(double),
                             hipMemcpyHostToDevice, s),
              "copy frequencies");
    auto lists = kern::build_brute_fold_lists(
        d_freqs.as<double>(), static_cast<uint32_t>(freqs.size()),
        static_cast<uint32_t>(segment_len), static_cast<uint32_t>(nbins), tsamp, t_ref, s);
    kern::brute_fold_time(ts_e, ts_v, nsamps / segment_len, lists, fold, FoldStorage::kF32, 1.0F,
                          s);
    check_launch("brute fold launch");
    // The lists are freed on return: finish before releasing them.
    hip_check(hipStreamSynchronize(s), "brute fold");
}

void brute_fold_fourier_device(const float* ts_e, const float* ts_v, SizeType nsamps,
                               std::span<const double> freqs, SizeType segment_len,
                               SizeType nbins, double tsamp, double t_ref, float2* fold,
                               hipStream_t s) {
    LOKI_REQUIRE(!freqs.empty(), "BruteFold: Frequency array is empty");
    LOKI_REQUIRE(segment_len > 0 && nsamps % segment_len == 0,
                 "BruteFold: Number of samples is not a multiple of segment length");
    LOKI_REQUIRE(nbins >= 2 && nbins <= 256, "BruteFold: Fourier fold needs nbins in [2, 256]");
    LOKI_REQUIRE(segment_len < (1ULL << 32) && freqs.size() < (1ULL << 32),
                 "BruteFold: segment too long");
    DeviceBuffer d_freqs(freqs.size() * sizeof(double));
    hip_check(hipMemcpyAsync(d_freqs.get(), freqs.data(), freqs.size() * sizeof(double),
                             hipMemcpyHostToDevice, s),
              "copy frequencies");
    kern::brute_fold_fourier(ts_e, ts_v, nsamps / segment_len,
                             static_cast<uint32_t>(segment_len), d_freqs.as<double>(),
                             static_cast<uint32_t>(freqs.size()), static_cast<uint32_t>(nbins),
                             tsamp, t_ref, fold, s);
    check_launch("Fourier brute fold launch");
    // The frequencies are freed on return: finish before releasing them.
    hip_check(hipStreamSynchronize(s), "Fourier brute fold");
}

}  // namespace loki_amd
