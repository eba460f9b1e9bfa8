// Host-side launchers of the gfx950 HIP kernels.  Every launcher is
// asynchronous on the given stream and validates shapes before launching.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "plan.hpp"

namespace loki_amd::kern {

/// Geometry of one merge level for the shift-table resolvers.  The last
/// three search parameters are (jerk, accel, freq); unused leading entries
/// have count 1.
struct ResolveLevel {
    uint32_t n_cur[3];
    uint32_t n_prev[3];
    double lo[3];
    double hi[3];
    double tseg_prev;  ///< tseg_brute * 2^(level-1)
    uint32_t nbins;
    uint32_t np;       ///< 1 (freq), 2 (accel, freq) or 3 (jerk, accel, freq)
};

/// Fills ncoords FFACoordFreq records (np == 1) of one level.
void resolve_freq(const ResolveLevel& lv, FFACoordFreq* out, uint64_t ncoords, hipStream_t s);
/// Fills ncoords FFACoord records (np == 2 or 3) of one level, tail and head.
void resolve_poly(const ResolveLevel& lv, FFACoord* out, uint64_t ncoords, hipStream_t s);

/// Bucket lists of the time-domain brute fold: for each frequency, the
/// sample indices of each phase bin in ascending order (CSR).
/// idx: [nfreqs * seglen] u32, off: [nfreqs * (nbins + 1)] u32.
void brute_fold_buckets(const double* freqs, uint32_t nfreqs, uint32_t seglen, uint32_t nbins,
                        double tsamp, double t_ref, uint32_t* idx, uint32_t* off,
                        hipStream_t s);
/// fold[nseg, nfreqs, 2, nbins] from the bucket lists (ascending-order sums).
void brute_fold_time(const float* ts_e, const float* ts_v, uint64_t nseg, uint32_t seglen,
                     uint32_t nfreqs, uint32_t nbins, const uint32_t* idx, const uint32_t* off,
                     float* fold, hipStream_t s);

/// Time-domain merge of one level (freq-only when coords_freq != nullptr).
void merge_time(const float* in, float* out, const FFACoord* coords,
                const FFACoordFreq* coords_freq, uint64_t ncur, uint64_t nprev, uint64_t nseg,
                uint32_t nbins, hipStream_t s);

/// Boxcar widths with the reference's template constants (score.cpp:107-116).
struct WidthTable {
    static constexpr uint32_t kMax = 64;
    uint32_t n;
    uint32_t wmax;
    uint32_t w[kMax];
    float h[kMax];
    float b[kMax];
};
WidthTable make_width_table(const SizeType* widths, SizeType nw, SizeType nbins);

/// snr_boxcar_3d[_max]: folds [nprof, 2, nbins] -> scores [nprof, nw] or [nprof].
void snr_boxcar_3d(const float* folds, uint64_t nprof, uint32_t nbins, const WidthTable& wt,
                   float* scores, bool find_max, hipStream_t s);
/// snr_boxcar_2d_max: folds [nprof, nbins] (already normalised) / stdnoise.
void snr_boxcar_2d_max(const float* folds, uint64_t nprof, uint32_t nbins, const WidthTable& wt,
                       float stdnoise, float* scores, hipStream_t s);

}  // namespace loki_amd::kern
