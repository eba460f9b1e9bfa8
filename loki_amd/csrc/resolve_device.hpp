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

/// Frequency of index jf of the last level's grid: the cell centre from the
/// precomputed dv_actual = (max - min) / count (dv), a single IEEE division
/// identical to psr_utils.hpp:290-292.
LOKI_HD inline double final_freq(const ResolveLevel& lv, const double* dv, uint32_t jf) {
    return (lv.n_cur[2] <= 1) ? (lv.hi[2] + lv.lo[2]) / 2.0
                              : lv.lo[2] + (dv[2] * (static_cast<double>(jf) + 0.5));
}

/// Row part of one side (0: tail, delta_t = -tseg/2; 1: head, +tseg/2) of
/// grid row `row` (flattened (jerk, accel) cell), NP == 2 or 3: the
/// taylor_ffa.cpp:19-188 terms that do not depend on the frequency, in the
/// reference's operation order.  Cell centres use dv as final_freq does.
template <int NP>
LOKI_HD inline RowTerms row_terms(const ResolveLevel& lv, const double* dv, uint32_t row,
                                  int side) {
    RowTerms t{};
    const uint32_t n_a   = lv.n_cur[1];
    const uint32_t ia    = row % n_a;
    const uint32_t ij    = row / n_a;
    const double a_cur   = (n_a <= 1) ? (lv.hi[1] + lv.lo[1]) / 2.0
                                      : lv.lo[1] + (dv[1] * (static_cast<double>(ia) + 0.5));
    const double delta_t = (static_cast<double>(side) - 0.5) * lv.tseg_prev;
    if constexpr (NP == 2) {
        (void)ij;
        const double half_delta_t_sq = 0.5 * delta_t * delta_t;
        const double v_new           = a_cur * delta_t;
        const double d_new           = a_cur * half_delta_t_sq;
        const uint32_t idx_a = grid::nearest_idx(a_cur, lv.lo[1], lv.hi[1], lv.n_prev[1]);
        t.fac  = 1.0 - (v_new * grid::kInvCval);
        t.td   = delta_t - (d_new * grid::kInvCval);
        t.base = idx_a * lv.n_prev[2];
    } else {
        const uint32_t n_j = lv.n_cur[0];
        const double j_cur = (n_j <= 1) ? (lv.hi[0] + lv.lo[0]) / 2.0
                                        : lv.lo[0] + (dv[0] * (static_cast<double>(ij) + 0.5));
        const double delta_t_sq            = delta_t * delta_t;
        const double delta_t_cubed         = delta_t_sq * delta_t;
        const double half_delta_t_sq       = 0.5 * delta_t_sq;
        const double sixth_delta_t_cubed   = delta_t_cubed / 6.0;
        const double j_delta_t             = j_cur * delta_t;
        const double half_j_delta_t_sq     = 0.5 * j_cur * delta_t_sq;
        const double j_sixth_delta_t_cubed = j_cur * sixth_delta_t_cubed;
        const uint32_t idx_j = grid::nearest_idx(j_cur, lv.lo[0], lv.hi[0], lv.n_prev[0]);
        const double a_new   = a_cur + j_delta_t;
        const double v_new   = (a_cur * delta_t) + half_j_delta_t_sq;
        const double d_new   = (a_cur * half_delta_t_sq) + j_sixth_delta_t_cubed;
        const uint32_t idx_a = grid::nearest_idx(a_new, lv.lo[1], lv.hi[1], lv.n_prev[1]);
        t.fac  = 1.0 - (v_new * grid::kInvCval);
        t.td   = delta_t - (d_new * grid::kInvCval);
        t.base = (idx_j * lv.n_prev[1] * lv.n_prev[2]) + (idx_a * lv.n_prev[2]);
    }
    return t;
}

/// Frequency part of one side: f_new = f_cur * (1 - v_new / c), the phase of
/// (delta_t - delay_rel) * f_cur and the parent index base + idx_f, the same
/// operations as the reference's per-coordinate expressions.
LOKI_HD inline void resolve_side_rt(const ResolveLevel& lv, const RowTerms& t, double f_cur,
                                    uint32_t& idx, float& shift) {
    const double f_new = f_cur * t.fac;
    shift              = grid::phase_idx_td(t.td, f_cur, lv.nbins);
    idx                = t.base + grid::nearest_idx(f_new, lv.lo[2], lv.hi[2], lv.n_prev[2]);
}

/// Both sides of the coordinate of grid row `row` at frequency index jf from
/// the row's terms (tail t0, head t1), NP == 2 or 3.
LOKI_HD inline FFACoord resolve_pair_rt(const ResolveLevel& lv, const double* dv,
                                        const RowTerms& t0, const RowTerms& t1, uint32_t jf) {
    FFACoord c;
    const double f_cur = final_freq(lv, dv, jf);
    resolve_side_rt(lv, t0, f_cur, c.i_tail, c.shift_tail);
    resolve_side_rt(lv, t1, f_cur, c.i_head, c.shift_head);
    return c;
}

/// Both sides of the coordinate in grid row `row` (flattened (jerk, accel)
/// cell) and frequency index jf; every value is produced by the same
/// operations as resolve_side (same bits).  Host and device.
template <int NP>
LOKI_HD inline FFACoord resolve_pair_rc(const ResolveLevel& lv, const double* dv, uint32_t row,
                                        uint32_t jf) {
    if constexpr (NP == 1) {
        (void)row;
        FFACoord c;
        const double f_cur = final_freq(lv, dv, jf);
        c.i_tail     = grid::nearest_idx(f_cur, lv.lo[2], lv.hi[2], lv.n_prev[2]);
        c.i_head     = c.i_tail;
        c.shift_tail = 0.0F;
        c.shift_head = grid::phase_idx(lv.tseg_prev, f_cur, lv.nbins, 0.0);
        return c;
    } else {
        return resolve_pair_rt(lv, dv, row_terms<NP>(lv, dv, row, 0),
                               row_terms<NP>(lv, dv, row, 1), jf);
    }
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
