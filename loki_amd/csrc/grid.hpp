// FP64 grid / phase helpers shared by host code and the HIP resolver kernels.
//
// Every expression keeps the reference's operation order and the library is
// compiled with -ffp-contract=off, so host and device produce identical bits.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define LOKI_HD __host__ __device__
#else
#define LOKI_HD
#endif

namespace loki_amd::grid {

// lib/loki/utils.hpp:11-14
inline constexpr double kCval     = 299792458.0;
inline constexpr double kInvCval  = 3.3356409519815204e-09;
inline constexpr double kFloatEps = 1e-6;

/// psr_utils.hpp:32-49 (phase_idx below) with its first factor
/// td = proper_time - delay formed by the caller: a grid row's resolver terms
/// hold it once per row.
LOKI_HD inline float phase_idx_td(double td, double freq, uint32_t nbins) {
    const double total = td * freq;
    const double norm  = total - floor(total);
    double iphase      = norm * static_cast<double>(nbins);
    if (iphase >= static_cast<double>(nbins)) {
        iphase = 0.0;
    }
    return static_cast<float>(iphase);
}

/// psr_utils.hpp:32-49: frac((t - delay) * f) * nbins, rounded to float;
/// a value equal to nbins wraps to 0.
LOKI_HD inline float phase_idx(double proper_time, double freq, uint32_t nbins,
                               double delay) {
    return phase_idx_td(proper_time - delay, freq, nbins);
}

/// psr_utils.hpp:53-63: integer phase bin u32(phase + 0.5f), nbins -> 0.
LOKI_HD inline uint32_t phase_bin(double proper_time, double freq, uint32_t nbins) {
    const float ph = phase_idx(proper_time, freq, nbins, 0.0);
    uint32_t b     = static_cast<uint32_t>(ph + 0.5F);
    if (b == nbins) {
        b = 0;
    }
    return b;
}

/// Unwrapped phase bin: floor(x) * nbins + w with x = t * freq and w the bin of
/// phase_bin before its nbins -> 0 wrap (w in [0, nbins]), computed with the
/// same FP64/FP32 operations.  phase_bin(t) == U mod nbins, and U is
/// non-decreasing in t (every step is a monotone rounding), so the samples of
/// one bin form contiguous runs, one per value of U.
LOKI_HD inline int64_t unwrapped_bin(double proper_time, double freq, uint32_t nbins) {
    const double total = (proper_time - 0.0) * freq;
    const double fl    = floor(total);
    const double norm  = total - fl;
    const double iphase = norm * static_cast<double>(nbins);
    uint32_t w = nbins;
    if (iphase < static_cast<double>(nbins)) {
        w = static_cast<uint32_t>(static_cast<float>(iphase) + 0.5F);
    }
    return static_cast<int64_t>(fl) * static_cast<int64_t>(nbins) + static_cast<int64_t>(w);
}

/// kernels.cpp:78-85 / kernels_cuda.cu:305-312: integer shift from a float
/// phase shift, u32(shift + 0.5f), nbins -> 0.
LOKI_HD inline uint32_t shift_bins(float shift, uint32_t nbins) {
    uint32_t s = static_cast<uint32_t>(shift + 0.5F);
    if (s == nbins) {
        s = 0;
    }
    return s;
}

/// psr_utils.hpp:283-293: cell centre vmin + dv_actual * (i + 0.5).
LOKI_HD inline double param_val_at_idx(double vmin, double vmax, uint32_t count,
                                       uint32_t i) {
    if (count <= 1) {
        return (vmax + vmin) / 2.0;
    }
    const double dv_actual = (vmax - vmin) / static_cast<double>(count);
    return vmin + (dv_actual * (static_cast<double>(i) + 0.5));
}

/// psr_utils.hpp:304-320: int(count*(val-min)/(max-min) + 1e-6), clamped.
LOKI_HD inline uint32_t nearest_idx(double val, double vmin, double vmax,
                                    uint32_t count) {
    if (count <= 1) {
        return 0;
    }
    const double raw = static_cast<double>(count) * (val - vmin) / (vmax - vmin);
    const double shifted = raw + kFloatEps;
    // Values far outside the grid clamp (the reference relies on small values).
    if (!(shifted > -1.0)) {
        return 0;
    }
    if (shifted >= static_cast<double>(count)) {
        return count - 1;
    }
    const int idx = static_cast<int>(shifted);
    if (idx < 0) {
        return 0;
    }
    return static_cast<uint32_t>(idx);
}

}  // namespace loki_amd::grid
