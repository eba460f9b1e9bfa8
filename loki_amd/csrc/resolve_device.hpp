// Device-side FP64 shift-table resolution of one coordinate (shared by the
// table kernels and the fused final-level kernel, so both produce the same
// bits).  Restates taylor_ffa.cpp:19-188 (accel / jerk, one side per call)
// and taylor_ffa.cpp:249-271 (freq-only) in the same operation order.
#pragma once

#include "grid.hpp"
#include "kernels.hpp"

namespace loki_amd::kern {

// One side (delta_t = -tseg/2: tail, +tseg/2: head) for NP == 2 (accel,
// freq) or NP == 3 (jerk, accel, freq).
template <int NP>
__device__ __forceinline__ void resolve_side(const ResolveLevel& lv, double delta_t,
                                             uint32_t ij, uint32_t ia, uint32_t jf,
                                             uint32_t& idx_out, float& shift_out) {
    const double f_cur = grid::param_val_at_idx(lv.lo[2], lv.hi[2], lv.n_cur[2], jf);
    const double a_cur = grid::param_val_at_idx(lv.lo[1], lv.hi[1], lv.n_cur[1], ia);
    if constexpr (NP == 2) {
        const double half_delta_t_sq = 0.5 * delta_t * delta_t;
        const double v_new           = a_cur * delta_t;
        const double d_new           = a_cur * half_delta_t_sq;
        const uint32_t idx_a = grid::nearest_idx(a_cur, lv.lo[1], lv.hi[1], lv.n_prev[1]);
        const double f_new     = f_cur * (1.0 - (v_new * grid::kInvCval));
        const double delay_rel = d_new * grid::kInvCval;
        shift_out = grid::phase_idx(delta_t, f_cur, lv.nbins, delay_rel);
        const uint32_t idx_f = grid::nearest_idx(f_new, lv.lo[2], lv.hi[2], lv.n_prev[2]);
        idx_out = idx_a * lv.n_prev[2] + idx_f;
    } else {
        const double delta_t_sq          = delta_t * delta_t;
        const double delta_t_cubed       = delta_t_sq * delta_t;
        const double half_delta_t_sq     = 0.5 * delta_t_sq;
        const double sixth_delta_t_cubed = delta_t_cubed / 6.0;
        const double j_cur = grid::param_val_at_idx(lv.lo[0], lv.hi[0], lv.n_cur[0], ij);
        const double j_delta_t             = j_cur * delta_t;
        const double half_j_delta_t_sq     = 0.5 * j_cur * delta_t_sq;
        const double j_sixth_delta_t_cubed = j_cur * sixth_delta_t_cubed;
        const uint32_t idx_j = grid::nearest_idx(j_cur, lv.lo[0], lv.hi[0], lv.n_prev[0]);
        const double a_new = a_cur + j_delta_t;
        const double v_new = (a_cur * delta_t) + half_j_delta_t_sq;
        const double d_new = (a_cur * half_delta_t_sq) + j_sixth_delta_t_cubed;
        const uint32_t idx_a = grid::nearest_idx(a_new, lv.lo[1], lv.hi[1], lv.n_prev[1]);
        const double f_new     = f_cur * (1.0 - (v_new * grid::kInvCval));
        const double delay_rel = d_new * grid::kInvCval;
        shift_out = grid::phase_idx(delta_t, f_cur, lv.nbins, delay_rel);
        const uint32_t idx_f = grid::nearest_idx(f_new, lv.lo[2], lv.hi[2], lv.n_prev[2]);
        idx_out = (idx_j * lv.n_prev[1] * lv.n_prev[2]) + (idx_a * lv.n_prev[2]) + idx_f;
    }
}

/// Both sides of the coordinate in grid row `row` (flattened (jerk, accel)
/// cell) and frequency index jf, with the per-coordinate terms evaluated once;
/// every value is produced by the same operations as resolve_side (same bits).
/// Cell centres use the precomputed dv_actual = (max - min) / count (dv), a
/// single IEEE division identical to psr_utils.hpp:290-292.  Host and device.
template <int NP>
LOKI_HD inline FFACoord resolve_pair_rc(const ResolveLevel& lv, const double* dv, uint32_t row,
                                        uint32_t jf) {
    FFACoord c;
    const uint32_t n_f = lv.n_cur[2];
    const double f_cur = (n_f <= 1) ? (lv.hi[2] + lv.lo[2]) / 2.0
                                    : lv.lo[2] + (dv[2] * (static_cast<double>(jf) + 0.5));
    if constexpr (NP == 1) {
        (void)row;
        c.i_tail     = grid::nearest_idx(f_cur, lv.lo[2], lv.hi[2], lv.n_prev[2]);
        c.i_head     = c.i_tail;
        c.shift_tail = 0.0F;
        c.shift_head = grid::phase_idx(lv.tseg_prev, f_cur, lv.nbins, 0.0);
    } else {
        const uint32_t n_a = lv.n_cur[1];
        const uint32_t ia  = row % n_a;
        const uint32_t ij  = row / n_a;
        const double a_cur = (n_a <= 1) ? (lv.hi[1] + lv.lo[1]) / 2.0
                                        : lv.lo[1] + (dv[1] * (static_cast<double>(ia) + 0.5));
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const double delta_t = (static_cast<double>(side) - 0.5) * lv.tseg_prev;
            uint32_t idx;
            float shift;
            if constexpr (NP == 2) {
                (void)ij;
                const double half_delta_t_sq = 0.5 * delta_t * delta_t;
                const double v_new           = a_cur * delta_t;
                const double d_new           = a_cur * half_delta_t_sq;
                const uint32_t idx_a = grid::nearest_idx(a_cur, lv.lo[1], lv.hi[1], lv.n_prev[1]);
                const double f_new     = f_cur * (1.0 - (v_new * grid::kInvCval));
                const double delay_rel = d_new * grid::kInvCval;
                shift = grid::phase_idx(delta_t, f_cur, lv.nbins, delay_rel);
                const uint32_t idx_f = grid::nearest_idx(f_new, lv.lo[2], lv.hi[2], lv.n_prev[2]);
                idx = idx_a * lv.n_prev[2] + idx_f;
            } else {
                const uint32_t n_j = lv.n_cur[0];
                const double j_cur = (n_j <= 1) ? (lv.hi[0] + lv.lo[0]) / 2.0
                                                : lv.lo[0] + (dv[0] * (static_cast<double>(ij) + 0.5));
                const double delta_t_sq          = delta_t * delta_t;
                const double delta_t_cubed       = delta_t_sq * delta_t;
                const double half_delta_t_sq     = 0.5 * delta_t_sq;
                const double sixth_delta_t_cubed = delta_t_cubed / 6.0;
                const double j_delta_t             = j_cur * delta_t;
                const double half_j_delta_t_sq     = 0.5 * j_cur * delta_t_sq;
                const double j_sixth_delta_t_cubed = j_cur * sixth_delta_t_cubed;
                const uint32_t idx_j = grid::nearest_idx(j_cur, lv.lo[0], lv.hi[0], lv.n_prev[0]);
                const double a_new = a_cur + j_delta_t;
                const double v_new = (a_cur * delta_t) + half_j_delta_t_sq;
                const double d_new = (a_cur * half_delta_t_sq) + j_sixth_delta_t_cubed;
                const uint32_t idx_a = grid::nearest_idx(a_new, lv.lo[1], lv.hi[1], lv.n_prev[1]);
                const double f_new     = f_cur * (1.0 - (v_new * grid::kInvCval));
                const double delay_rel = d_new * grid::kInvCval;
                shift = grid::phase_idx(delta_t, f_cur, lv.nbins, delay_rel);
                const uint32_t idx_f = grid::nearest_idx(f_new, lv.lo[2], lv.hi[2], lv.n_prev[2]);
                idx = (idx_j * lv.n_prev[1] * lv.n_prev[2]) + (idx_a * lv.n_prev[2]) + idx_f;
            }
            if (side == 0) {
                c.i_tail     = idx;
                c.shift_tail = shift;
            } else {
                c.i_head     = idx;
                c.shift_head = shift;
            }
        }
    }
    return c;
}

/// resolve_pair_rc of flat coordinate i = row * n_f + jf.
template <int NP>
__device__ __forceinline__ FFACoord resolve_pair(const ResolveLevel& lv, const double* dv,
                                                 uint64_t i) {
    const uint32_t n_f = lv.n_cur[2];
    return resolve_pair_rc<NP>(lv, dv, static_cast<uint32_t>(i / n_f),
                               static_cast<uint32_t>(i % n_f));
}

/// Full merge record of coordinate i of a level (NP == 1: freq-only, stored
/// as i_tail == i_head == idx, shift_tail == 0).
template <int NP>
__device__ __forceinline__ FFACoord resolve_coord(const ResolveLevel& lv, uint64_t i) {
    FFACoord c;
    if constexpr (NP == 1) {
        const double f_cur = grid::param_val_at_idx(lv.lo[2], lv.hi[2], lv.n_cur[2],
                                                    static_cast<uint32_t>(i));
        c.i_tail     = grid::nearest_idx(f_cur, lv.lo[2], lv.hi[2], lv.n_prev[2]);
        c.i_head     = c.i_tail;
        c.shift_tail = 0.0F;
        c.shift_head = grid::phase_idx(lv.tseg_prev, f_cur, lv.nbins, 0.0);
    } else {
        // Decompose (f fastest): i = (j * n_a + a) * n_f + f
        const uint32_t n_f = lv.n_cur[2];
        const uint32_t n_a = lv.n_cur[1];
        const uint32_t jf  = static_cast<uint32_t>(i % n_f);
        const uint64_t r   = i / n_f;
        const uint32_t ia  = static_cast<uint32_t>(r % n_a);
        const uint32_t ij  = static_cast<uint32_t>(r / n_a);
        resolve_side<NP>(lv, (0.0 - 0.5) * lv.tseg_prev, ij, ia, jf, c.i_tail, c.shift_tail);
        resolve_side<NP>(lv, (1.0 - 0.5) * lv.tseg_prev, ij, ia, jf, c.i_head, c.shift_head);
    }
    return c;
}

}  // namespace loki_amd::kern
