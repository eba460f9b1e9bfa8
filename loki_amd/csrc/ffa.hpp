// FFA executor on one GPU: the counterpart of loki's FFACUDA<T>
// (include/loki/algorithms/ffa.hpp:100-153, lib/ffa_cuda.cu:28-524) and of
// FFAWorkspaceCUDA (workspace.hpp:169-206).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <span>
#include <vector>

#include "plan.hpp"

namespace loki_amd {

/// RAII device allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(SizeType bytes);
    ~DeviceBuffer();
    DeviceBuffer(DeviceBuffer&& o) noexcept : m_ptr(o.m_ptr), m_bytes(o.m_bytes) {
        o.m_ptr   = nullptr;
        o.m_bytes = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    DeviceBuffer(const DeviceBuffer&)            = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void* get() const noexcept { return m_ptr; }
    SizeType bytes() const noexcept { return m_bytes; }
    template <typename T> T* as() const noexcept { return static_cast<T*>(m_ptr); }
    /// Grow (never shrink) to at least `bytes`; contents are not preserved.
    void reserve(SizeType bytes);

private:
    void* m_ptr{nullptr};
    SizeType m_bytes{0};
};

void hip_check(hipError_t err, const char* what);

/// Current-device guard (cuda_utils.cuh:434 CudaSetDeviceGuard).
class DeviceGuard {
public:
    explicit DeviceGuard(int device);
    ~DeviceGuard();
    DeviceGuard(const DeviceGuard&)            = delete;
    DeviceGuard& operator=(const DeviceGuard&) = delete;

private:
    int m_prev{0};
    bool m_changed{false};
};

class FFAExecutor {
public:
    FFAExecutor(const SearchConfig& cfg, FoldKind kind, int device_id);

    const FFAPlan& plan() const noexcept { return m_plan; }
    int device_id() const noexcept { return m_device; }

    /// Device execute (ffa_cuda.cu:171-196): ts_e/ts_v have nsamps floats;
    /// fold has plan().buffer_size() elements (float, or float2 for Fourier);
    /// the result is the first fold_size() elements.  Asynchronous on `s`.
    void execute(const float* ts_e, const float* ts_v, void* fold, SizeType fold_len,
                 hipStream_t s);

    /// Time-domain FFA + boxcar S/N of the final level (nsegments must be 1):
    /// scores [ncoords_final, nw] (or [ncoords_final] when find_max).
    void execute_scores(const float* ts_e, const float* ts_v, std::span<const SizeType> widths,
                        float* scores, SizeType scores_len, bool find_max, hipStream_t s);

    /// Resolve all levels' shift tables into device memory (level 0 zeroed).
    void resolve_coordinates(void* coords_dev, hipStream_t s) const;

    /// Device pointer to the last shift tables resolved by execute().
    const void* coords_device() const noexcept { return m_coords.get(); }
    SizeType coord_record_bytes() const noexcept;

private:
    FFAPlan m_plan;
    int m_device;
    bool m_freq_only;
    DeviceBuffer m_internal;   // ping-pong partner (buffer_size elements)
    DeviceBuffer m_result;     // second buffer for execute_scores
    DeviceBuffer m_coords;     // all levels' shift tables
    DeviceBuffer m_bf_freqs;   // level-0 frequencies (FP64)
    DeviceBuffer m_bf_idx;     // bucket lists
    DeviceBuffer m_bf_off;
    bool m_buckets_ready{false};

    void prepare_brute_fold(hipStream_t s);
    void run_levels(const float* ts_e, const float* ts_v, void* out_final, void* other,
                    hipStream_t s);
};

/// Standalone brute fold (fold_cuda.cu:139-164, 260-277): fold has
/// nsamps/segment_len * nfreqs * 2 * nbins floats.
void brute_fold_time_device(const float* ts_e, const float* ts_v, SizeType nsamps,
                            std::span<const double> freqs, SizeType segment_len, SizeType nbins,
                            double tsamp, double t_ref, float* fold, hipStream_t s);

}  // namespace loki_amd
