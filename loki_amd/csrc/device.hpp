// Device memory / device selection helpers (HIP runtime).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

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

}  // namespace loki_amd
