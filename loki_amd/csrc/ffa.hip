// FFA level driver on one GPU (restates ffa_cuda.cu:171-196, 409-522 and
// workspace_cuda.cu:83-158 without their per-level host synchronisations):
// shift tables for every level, the brute fold and all merge levels are
// enqueued back to back on the caller's stream.
#include "ffa.hpp"

#include <cmath>
#include <string>

#include "kernels.hpp"

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

namespace {

kern::ResolveLevel make_level(const FFAPlan& plan, SizeType l) {
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

void check_launch(const char* what) { hip_check(hipGetLastError(), what); }

}  // namespace

FFAExecutor::FFAExecutor(const SearchConfig& cfg, FoldKind kind, int device_id)
    : m_plan(cfg, kind), m_device(device_id), m_freq_only(cfg.get_nparams() == 1) {
    DeviceGuard g(device_id);
    LOKI_REQUIRE(kind == FoldKind::kTime,
                 "the Fourier-domain FFA is not available in this build");
    LOKI_REQUIRE(cfg.get_nbins() <= 1024, "nbins must be <= 1024");
    for (SizeType l = 1; l < m_plan.n_levels(); ++l) {
        LOKI_REQUIRE(m_plan.ncoords()[l - 1] < (1ULL << 32),
                     "parent grid exceeds 2^32 coordinates (u32 indices)");
    }
}

SizeType FFAExecutor::coord_record_bytes() const noexcept {
    return m_freq_only ? sizeof(FFACoordFreq) : sizeof(FFACoord);
}

void FFAExecutor::resolve_coordinates(void* coords_dev, hipStream_t s) const {
    auto* base = static_cast<char*>(coords_dev);
    for (SizeType l = 1; l < m_plan.n_levels(); ++l) {
        const auto lv      = make_level(m_plan, l);
        const SizeType off = m_plan.ncoords_offsets64()[l];
        const SizeType n   = m_plan.ncoords()[l];
        if (m_freq_only) {
            kern::resolve_freq(lv, reinterpret_cast<FFACoordFreq*>(base) + off, n, s);
        } else {
            kern::resolve_poly(lv, reinterpret_cast<FFACoord*>(base) + off, n, s);
        }
        check_launch("resolver launch");
    }
}

void FFAExecutor::prepare_brute_fold(hipStream_t s) {
    if (m_buckets_ready) {
        return;
    }
    const auto freqs       = m_plan.compute_param_grid(0).back();
    const SizeType nfreqs  = freqs.size();
    const SizeType seglen  = m_plan.segment_lens()[0];
    const SizeType nbins   = m_plan.config().get_nbins();
    const double t_ref     = m_freq_only ? 0.0 : m_plan.tsegments()[0] / 2.0;
    LOKI_REQUIRE(seglen < (1ULL << 32) && nfreqs < (1ULL << 32), "brute fold too large");
    m_bf_freqs.reserve(nfreqs * sizeof(double));
    m_bf_idx.reserve(nfreqs * seglen * sizeof(uint32_t));
    m_bf_off.reserve(nfreqs * (nbins + 1) * sizeof(uint32_t));
    hip_check(hipMemcpyAsync(m_bf_freqs.get(), freqs.data(), nfreqs * sizeof(double),
                             hipMemcpyHostToDevice, s),
              "copy brute-fold frequencies");
    kern::brute_fold_buckets(m_bf_freqs.as<double>(), static_cast<uint32_t>(nfreqs),
                             static_cast<uint32_t>(seglen), static_cast<uint32_t>(nbins),
                             m_plan.config().get_tsamp(), t_ref, m_bf_idx.as<uint32_t>(),
                             m_bf_off.as<uint32_t>(), s);
    check_launch("brute-fold bucket launch");
    // The host vector `freqs` must outlive the async copy.
    hip_check(hipStreamSynchronize(s), "brute-fold setup");
    m_buckets_ready = true;
}

void FFAExecutor::run_levels(const float* ts_e, const float* ts_v, void* out_final,
                             void* other, hipStream_t s) {
    const SizeType L = m_plan.n_levels();
    m_coords.reserve(m_plan.coord_size() * coord_record_bytes());
    prepare_brute_fold(s);
    resolve_coordinates(m_coords.get(), s);

    // ffa.cpp:261-276: start on the side that makes the last level land in
    // out_final.
    const bool odd_swaps = ((L - 2) % 2) == 1;
    auto* cur_in         = static_cast<float*>(odd_swaps ? out_final : other);
    auto* cur_out        = static_cast<float*>(odd_swaps ? other : out_final);

    const SizeType nbins = m_plan.config().get_nbins();
    kern::brute_fold_time(ts_e, ts_v, m_plan.nsegments()[0],
                          static_cast<uint32_t>(m_plan.segment_lens()[0]),
                          static_cast<uint32_t>(m_plan.ncoords()[0]),
                          static_cast<uint32_t>(nbins), m_bf_idx.as<uint32_t>(),
                          m_bf_off.as<uint32_t>(), cur_in, s);
    check_launch("brute fold launch");

    for (SizeType l = 1; l < L; ++l) {
        const SizeType off = m_plan.ncoords_offsets64()[l];
        const FFACoord* cp = nullptr;
        const FFACoordFreq* cfp = nullptr;
        if (m_freq_only) {
            cfp = m_coords.as<FFACoordFreq>() + off;
        } else {
            cp = m_coords.as<FFACoord>() + off;
        }
        kern::merge_time(cur_in, cur_out, cp, cfp, m_plan.ncoords()[l], m_plan.ncoords()[l - 1],
                         m_plan.nsegments()[l], static_cast<uint32_t>(nbins), s);
        check_launch("merge launch");
        if (l + 1 < L) {
            std::swap(cur_in, cur_out);
        }
    }
}

void FFAExecutor::execute(const float* ts_e, const float* ts_v, void* fold, SizeType fold_len,
                          hipStream_t s) {
    DeviceGuard g(m_device);
    LOKI_REQUIRE(ts_e != nullptr && ts_v != nullptr && fold != nullptr,
                 "FFA::execute: null buffer");
    LOKI_REQUIRE(fold_len == m_plan.buffer_size(),
                 "FFA::execute: fold must have size buffer_size");
    const SizeType elem = sizeof(float);
    m_internal.reserve(m_plan.buffer_size() * elem);
    run_levels(ts_e, ts_v, fold, m_internal.get(), s);
}

void FFAExecutor::execute_scores(const float* ts_e, const float* ts_v,
                                 std::span<const SizeType> widths, float* scores,
                                 SizeType scores_len, bool find_max, hipStream_t s) {
    DeviceGuard g(m_device);
    LOKI_REQUIRE(m_plan.nsegments().back() == 1,
                 "compute_ffa_scores: nsegments must be 1 for scores");
    const SizeType nbins   = m_plan.config().get_nbins();
    const SizeType ncoords = m_plan.ncoords().back();
    const auto wt          = kern::make_width_table(widths.data(), widths.size(), nbins);
    LOKI_REQUIRE(scores != nullptr, "scores buffer is null");
    LOKI_REQUIRE(scores_len == (find_max ? ncoords : ncoords * widths.size()),
                 "scores must have size ncoords * nwidths (or ncoords for the max)");
    m_internal.reserve(m_plan.buffer_size() * sizeof(float));
    m_result.reserve(m_plan.buffer_size() * sizeof(float));
    run_levels(ts_e, ts_v, m_result.get(), m_internal.get(), s);
    kern::snr_boxcar_3d(m_result.as<float>(), ncoords, static_cast<uint32_t>(nbins), wt, scores,
                        find_max, s);
    check_launch("score launch");
}

void brute_fold_time_device(const float* ts_e, const float* ts_v, SizeType nsamps,
                            std::span<const double> freqs, SizeType segment_len, SizeType nbins,
                            double tsamp, double t_ref, float* fold, hipStream_t s) {
    LOKI_REQUIRE(!freqs.empty(), "BruteFold: Frequency array is empty");
    LOKI_REQUIRE(segment_len > 0 && nsamps % segment_len == 0,
                 "BruteFold: Number of samples is not a multiple of segment length");
    LOKI_REQUIRE(nbins >= 1 && nbins <= 1024, "BruteFold: nbins must be in [1, 1024]");
    DeviceBuffer d_freqs(freqs.size() * sizeof(double));
    DeviceBuffer d_idx(freqs.size() * segment_len * sizeof(uint32_t));
    DeviceBuffer d_off(freqs.size() * (nbins + 1) * sizeof(uint32_t));
    hip_check(hipMemcpyAsync(d_freqs.get(), freqs.data(), freqs.size() * sizeof(double),
                             hipMemcpyHostToDevice, s),
              "copy frequencies");
    kern::brute_fold_buckets(d_freqs.as<double>(), static_cast<uint32_t>(freqs.size()),
                             static_cast<uint32_t>(segment_len), static_cast<uint32_t>(nbins),
                             tsamp, t_ref, d_idx.as<uint32_t>(), d_off.as<uint32_t>(), s);
    check_launch("brute-fold bucket launch");
    kern::brute_fold_time(ts_e, ts_v, nsamps / segment_len, static_cast<uint32_t>(segment_len),
                          static_cast<uint32_t>(freqs.size()), static_cast<uint32_t>(nbins),
                          d_idx.as<uint32_t>(), d_off.as<uint32_t>(), fold, s);
    check_launch("brute fold launch");
    // Temporaries are freed on return: finish before releasing them.
    hip_check(hipStreamSynchronize(s), "brute fold");
}

}  // namespace loki_amd
