// FP64 shift-table resolvers: one thread per coordinate of one merge level.
//
// Restates taylor_ffa.cpp:249-271 (freq-only) and taylor_ffa.cpp:19-188
// (accel / jerk, tail and head) with the same operation order; compiled with
// -ffp-contract=off so the tables are bit-identical to the FP64 oracle.
#include "grid.hpp"
#include "kernels.hpp"

namespace loki_amd::kern {
namespace {

__global__ void __launch_bounds__(256)
k_resolve_freq(ResolveLevel lv, FFACoordFreq* __restrict__ out, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double f_cur = grid::param_val_at_idx(lv.lo[2], lv.hi[2], lv.n_cur[2],
                                                static_cast<uint32_t>(i));
    FFACoordFreq c;
    c.idx   = grid::nearest_idx(f_cur, lv.lo[2], lv.hi[2], lv.n_prev[2]);
    c.shift = grid::phase_idx(lv.tseg_prev, f_cur, lv.nbins, 0.0);
    out[i]  = c;
}

// One side (LATTER 0 = tail, 1 = head) of taylor_ffa.cpp:19-86 (NP == 2) or
// taylor_ffa.cpp:88-188 (NP == 3).
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

template <int NP>
__global__ void __launch_bounds__(256)
k_resolve_poly(ResolveLevel lv, FFACoord* __restrict__ out, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    // Decompose (f fastest): i = (j * n_a + a) * n_f + f
    const uint32_t n_f = lv.n_cur[2];
    const uint32_t n_a = lv.n_cur[1];
    const uint32_t jf  = static_cast<uint32_t>(i % n_f);
    const uint64_t r   = i / n_f;
    const uint32_t ia  = static_cast<uint32_t>(r % n_a);
    const uint32_t ij  = static_cast<uint32_t>(r / n_a);
    FFACoord c;
    resolve_side<NP>(lv, (0.0 - 0.5) * lv.tseg_prev, ij, ia, jf, c.i_tail, c.shift_tail);
    resolve_side<NP>(lv, (1.0 - 0.5) * lv.tseg_prev, ij, ia, jf, c.i_head, c.shift_head);
    out[i] = c;
}

unsigned grid_for(uint64_t n, unsigned block) {
    const uint64_t g = (n + block - 1) / block;
    LOKI_REQUIRE(g < (1ULL << 31), "resolver grid too large");
    return static_cast<unsigned>(g);
}

}  // namespace

void resolve_freq(const ResolveLevel& lv, FFACoordFreq* out, uint64_t ncoords, hipStream_t s) {
    LOKI_REQUIRE(lv.np == 1, "resolve_freq needs np == 1");
    LOKI_REQUIRE(ncoords == lv.n_cur[2], "resolve_freq: coords size mismatch");
    if (ncoords == 0) {
        return;
    }
    k_resolve_freq<<<grid_for(ncoords, 256), 256, 0, s>>>(lv, out, ncoords);
}

void resolve_poly(const ResolveLevel& lv, FFACoord* out, uint64_t ncoords, hipStream_t s) {
    LOKI_REQUIRE(lv.np == 2 || lv.np == 3, "resolve_poly needs np in {2, 3}");
    LOKI_REQUIRE(ncoords == static_cast<uint64_t>(lv.n_cur[0]) * lv.n_cur[1] * lv.n_cur[2],
                 "resolve_poly: coords size mismatch");
    if (ncoords == 0) {
        return;
    }
    if (lv.np == 2) {
        k_resolve_poly<2><<<grid_for(ncoords, 256), 256, 0, s>>>(lv, out, ncoords);
    } else {
        k_resolve_poly<3><<<grid_for(ncoords, 256), 256, 0, s>>>(lv, out, ncoords);
    }
}

}  // namespace loki_amd::kern
