// FP64 shift-table resolvers: one thread per coordinate of one merge level.
//
// Restates taylor_ffa.cpp:249-271 (freq-only) and taylor_ffa.cpp:19-188
// (accel / jerk, tail and head) with the same operation order; compiled with
// -ffp-contract=off so the tables are bit-identical to the FP64 oracle.
#include "resolve_device.hpp"

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

template <int NP>
__global__ void __launch_bounds__(256)
k_resolve_poly(ResolveLevel lv, FFACoord* __restrict__ out, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const FFACoord c = resolve_coord<NP>(lv, i);
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
