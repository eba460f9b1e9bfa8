// FP64 shift-table resolvers: one thread per coordinate of one merge level.
//
// Restates taylor_ffa.cpp:249-271 (freq-only) and taylor_ffa.cpp:19-188
// (accel / jerk, tail and head) with the same operation order; compiled with
// -ffp-contract=off so the tables are bit-identical to the FP64 oracle.  The
// accel / jerk tables take each grid row's frequency-independent terms from a
// per-level table formed once on the host by the same host/device code.
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

// One workgroup per (grid row, 256 frequencies): the row's terms are
// workgroup-uniform scalar loads and each lane resolves one frequency.
__global__ void __launch_bounds__(256)
k_resolve_poly(ResolveLevel lv, FinalGrid fg, const RowTerms* __restrict__ rows,
               FFACoord* __restrict__ out, uint32_t tiles_per_row) {
    const uint32_t row  = blockIdx.x / tiles_per_row;
    const uint32_t tile = blockIdx.x - row * tiles_per_row;
    const uint32_t jf   = tile * 256 + threadIdx.x;
    const uint32_t n_f  = lv.n_cur[2];
    if (jf >= n_f) {
        return;
    }
    out[static_cast<uint64_t>(row) * n_f + jf] =
        resolve_pair_rt(lv, fg.dv, rows[2 * static_cast<size_t>(row)],
                        rows[2 * static_cast<size_t>(row) + 1], jf);
}

__device__ __forceinline__ uint32_t find_task_resolve(const ResolveTask* __restrict__ tasks,
                                                      uint32_t ntasks) {
    uint32_t lo = 0;
    uint32_t hi = ntasks;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (tasks[mid].block0 <= blockIdx.x) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    return lo;
}

template <int NP>
__global__ void __launch_bounds__(256)
k_resolve_grouped(const ResolveTask* __restrict__ tasks, uint32_t ntasks) {
    const ResolveTask& t = tasks[find_task_resolve(tasks, ntasks)];
    const uint64_t i     = static_cast<uint64_t>(blockIdx.x - t.block0) * 256 + threadIdx.x;
    if (i >= t.n) {
        return;
    }
    if constexpr (NP == 1) {
        const double f_cur = grid::param_val_at_idx(t.lv.lo[2], t.lv.hi[2], t.lv.n_cur[2],
                                                    static_cast<uint32_t>(i));
        FFACoordFreq c;
        c.idx = grid::nearest_idx(f_cur, t.lv.lo[2], t.lv.hi[2], t.lv.n_prev[2]);
        c.shift = grid::phase_idx(t.lv.tseg_prev, f_cur, t.lv.nbins, 0.0);
        static_cast<FFACoordFreq*>(t.out)[i] = c;
    } else {
        static_cast<FFACoord*>(t.out)[i] = resolve_coord<NP>(t.lv, i);
    }
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

void resolve_poly(const ResolveLevel& lv, const RowTerms* rows, FFACoord* out, uint64_t ncoords,
                  hipStream_t s) {
    LOKI_REQUIRE(lv.np == 2 || lv.np == 3, "resolve_poly needs np in {2, 3}");
    LOKI_REQUIRE(ncoords == static_cast<uint64_t>(lv.n_cur[0]) * lv.n_cur[1] * lv.n_cur[2],
                 "resolve_poly: coords size mismatch");
    LOKI_REQUIRE(rows != nullptr, "resolve_poly: row terms missing");
    if (ncoords == 0) {
        return;
    }
    FinalGrid fg{};
    for (int k = 0; k < 3; ++k) {
        fg.dv[k] = (lv.hi[k] - lv.lo[k]) / static_cast<double>(lv.n_cur[k]);
    }
    const uint64_t nrows = static_cast<uint64_t>(lv.n_cur[0]) * lv.n_cur[1];
    const uint32_t tpr   = (lv.n_cur[2] + 255) / 256;
    const unsigned grid  = grid_for(nrows * tpr * 256, 256);
    k_resolve_poly<<<grid, 256, 0, s>>>(lv, fg, rows, out, tpr);
}

void resolve_grouped(const ResolveTask* tasks, uint32_t ntasks, uint32_t nblocks, uint32_t np,
                     hipStream_t s) {
    if (ntasks == 0 || nblocks == 0) {
        return;
    }
    switch (np) {
    case 1: k_resolve_grouped<1><<<nblocks, 256, 0, s>>>(tasks, ntasks); break;
    case 2: k_resolve_grouped<2><<<nblocks, 256, 0, s>>>(tasks, ntasks); break;
    case 3: k_resolve_grouped<3><<<nblocks, 256, 0, s>>>(tasks, ntasks); break;
    default: LOKI_REQUIRE(false, "resolve_grouped: np must be 1, 2 or 3");
    }
}

}  // namespace loki_amd::kern
