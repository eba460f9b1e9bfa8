// Persistent batch kernel: every stage of every series of a batch -- shift
// tables, brute folds, all merge levels and the fused last levels -- in one
// launch.
//
// The host lays the batch out as items (one stage of one series), each a
// contiguous range of tickets, one ticket per tile of 256 lanes, with the
// items sorted so that every item's dependencies (the previous level and the
// level's shift table of the same series) hold smaller tickets.  Resident
// workgroups take tickets in order from one counter.  Before its tile a
// workgroup waits until each dependency's tile counter reaches the
// dependency's tile count (agent-scope acquire: the other XCDs' L2 lines of
// those rows are not trusted), and after it the workgroup adds one to its
// own item's counter (every lane fences its stores first, agent-scope
// release).  A workgroup only ever waits on smaller tickets, all of them
// taken by running workgroups, so the smallest unfinished ticket always
// proceeds and the launch drains; a bounded wait (spin_limit polls) raises an
// error flag instead of hanging should that ever fail.  The last workgroup to
// finish zeroes the counters for the next launch, so a captured graph replays
// with no memset node.
//
// Each tile runs the same per-tile code as the grouped kernels (fold_device,
// merge_device, final_device), so every series' scores are bit-identical to
// the grouped launches and to the single-series path.  The fused last level
// uses all 256 lanes: a tile holds 256 / WS trials and WS lanes score one
// trial, lane q taking widths q, q + WS, ... (the boxcar-width schedule the
// batch autotuner chooses per nbins).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "final_device.hpp"
#include "fold_device.hpp"
#include "kernels.hpp"
#include "merge_device.hpp"
#include "resolve_device.hpp"

namespace loki_amd::kern {
namespace {

static_assert(kTreeBlock == static_cast<uint32_t>(kFoldBlock) && kTreeBlock == kFinalWideBlock &&
                  kTreeBlock == static_cast<uint32_t>(kMergeBlock),
              "tiles of every stage are 256 lanes");

/// Shift-table records of tile `tile` of a freq-only level (the records of
/// k_resolve_grouped<1>).
__device__ __forceinline__ void resolve_tile(uint32_t tile, const ResolveTask& t) {
    const uint64_t i = static_cast<uint64_t>(tile) * kTreeBlock + threadIdx.x;
    if (i < t.n) {
        const FFACoord c = resolve_coord<1>(t.lv, i);
        FFACoordFreq r;
        r.idx   = c.i_tail;
        r.shift = c.shift_head;
        static_cast<FFACoordFreq*>(t.out)[i] = r;
    }
}

template <int NB>
__device__ __forceinline__ void merge_tile_nb(uint32_t tile, const MergeTask& t) {
    merge_v4_block<NB, false, V4Geom<NB>::ITERS, float, float>(
        tile, static_cast<const float*>(t.in), static_cast<float*>(t.out), nullptr,
        t.coords_freq, t.ncur, t.nprev, t.ts, t.scale, nullptr, nullptr, false);
}

template <int NB>
__device__ __forceinline__ void final_tile_ws(uint32_t lanes, uint32_t tile, const FinalTask& t,
                                              unsigned char* lds) {
    if constexpr (NB >= 16) {
        if (lanes == 4) {
            return final_tile<NB, 4>(tile, t, lds);
        }
    }
    if constexpr (NB <= 64) {
        if (lanes == 2) {
            return final_tile<NB, 2>(tile, t, lds);
        }
    }
    if constexpr (NB <= 32) {
        final_tile<NB, 1>(tile, t, lds);
    }
}

/// Brute-fold tiles (the grouped kernels' block code).
__device__ __forceinline__ void fold_list_tile(uint32_t tile, const FoldTask& f,
                                                         unsigned char* lds) {
    auto* s_e = reinterpret_cast<float*>(lds);
    bf_fold_block<true, float>(tile, f.ts_e, f.ts_v, f.seglen, f.nfreqs, f.nbins, f.bps, f.list,
                               f.base, f.counts, static_cast<float*>(f.fold), f.scale, s_e,
                               s_e + kLdsSamps);
}

__device__ __forceinline__ void fold_units_tile(uint32_t tile, const FoldTask& f) {
    bf_units_block<float>(tile, f.ts_e, f.ts_v, f.seglen, f.nfreqs, f.nbins, f.bps, f.freqs,
                          f.tsamp, f.t_ref, static_cast<float*>(f.fold), f.scale);
}

/// Waits (lane 0) until every tile of item `dep` is done.  The counter is
/// polled with relaxed agent-scope loads (no cache maintenance per poll: an
/// acquire on every poll would invalidate the XCD's L2 under the workgroups
/// that are working) and the L2 is invalidated once, by the caller's acquire
/// fence, when the count is reached.  False if the wait gave up: after
/// spin_limit polls, or at once when another wait of the launch already did
/// (so a failed launch still ends after one time-out).
__device__ __forceinline__ bool wait_item(const TreeArgs& a, uint32_t dep) {
    if (dep == kTreeNoDep) {
        return true;
    }
    const uint32_t need = a.items[dep].ntiles;
    uint32_t* ctr       = a.sync + kTreeSyncHeader + dep;
    for (uint32_t polls = 0;; ++polls) {
        if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need) {
            return true;
        }
        if (polls >= a.spin_limit ||
            __hip_atomic_load(a.sync + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            return false;
        }
        __builtin_amdgcn_s_sleep(32);
    }
}

__global__ void __launch_bounds__(kTreeBlock) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_tree(TreeArgs a) {
    extern __shared__ __align__(16) unsigned char s_lds[];
    __shared__ uint32_t s_ticket;
    __shared__ uint32_t s_last;
    const uint32_t tid = threadIdx.x;
    for (;;) {
        if (tid == 0) {
            s_ticket = __hip_atomic_fetch_add(a.sync, 1U, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const uint32_t ticket = s_ticket;
        if (ticket >= a.ntickets) {
            break;
        }
        uint32_t lo = 0;  // the item holding the ticket
        uint32_t hi = a.nitems;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (a.items[mid].ticket0 <= ticket) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        const TreeItem it = a.items[lo];
        if (tid == 0 && (it.dep[0] != kTreeNoDep || it.dep[1] != kTreeNoDep)) {
            const bool ok = wait_item(a, it.dep[0]) && wait_item(a, it.dep[1]);
            if (!ok) {
                __hip_atomic_fetch_or(a.sync + 2, 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            // acquire: the dependencies' rows are read past any stale L1/L2 lines
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        const uint32_t tile = ticket - it.ticket0;
        switch (static_cast<TreeKind>(it.kind)) {
        case TreeKind::kResolve: resolve_tile(tile, a.resolve[it.task]); break;
        case TreeKind::kFoldList: fold_list_tile(tile, a.fold[it.task], s_lds); break;
        case TreeKind::kFoldUnits: fold_units_tile(tile, a.fold[it.task]); break;
        case TreeKind::kMerge: {
            const MergeTask& m = a.merge[it.task];
            switch (it.nbins) {
            case 8: merge_tile_nb<8>(tile, m); break;
            case 16: merge_tile_nb<16>(tile, m); break;
            case 32: merge_tile_nb<32>(tile, m); break;
            case 64: merge_tile_nb<64>(tile, m); break;
            case 128: merge_tile_nb<128>(tile, m); break;
            default: break;
            }
            break;
        }
        case TreeKind::kFinal: {
            const FinalTask& f = a.final_[it.task];
            switch (it.nbins) {
            case 8: final_tile_ws<8>(it.lanes, tile, f, s_lds); break;
            case 16: final_tile_ws<16>(it.lanes, tile, f, s_lds); break;
            case 32: final_tile_ws<32>(it.lanes, tile, f, s_lds); break;
            case 64: final_tile_ws<64>(it.lanes, tile, f, s_lds); break;
            case 128: final_tile_ws<128>(it.lanes, tile, f, s_lds); break;
            default: break;
            }
            break;
        }
        }
        // the barrier completes every wave's stores (workgroup-scope release)
        // and frees s_ticket and the LDS for the next tile; lane 0's
        // agent-scope release then writes the XCD's L2 back before the count
        __syncthreads();
        if (tid == 0) {
            __hip_atomic_fetch_add(a.sync + kTreeSyncHeader + lo, 1U, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // the last workgroup out zeroes the counters for the next launch
    if (tid == 0) {
        const uint32_t done =
            __hip_atomic_fetch_add(a.sync + 1, 1U, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (done + 1 == gridDim.x) ? 1U : 0U;
    }
    __syncthreads();
    if (s_last != 0) {
        for (uint32_t i = tid; i < a.nitems; i += kTreeBlock) {
            a.sync[kTreeSyncHeader + i] = 0;
        }
        if (tid == 0) {
            a.sync[0] = 0;
            a.sync[1] = 0;
        }
    }
}

}  // namespace

size_t tree_lds_bytes(TreeKind kind, uint32_t nbins, uint32_t lanes, uint32_t stride) {
    switch (kind) {
    case TreeKind::kFoldList: return 2 * static_cast<size_t>(kLdsSamps) * sizeof(float);
    case TreeKind::kFoldUnits: return unit_lds_bytes(nbins);
    case TreeKind::kFinal: return final_wide_lds_bytes(lanes, stride);
    default: return 0;
    }
}

uint32_t tree_resident_blocks(size_t lds) {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    int ncu = 0;
    hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev),
              "hipDeviceGetAttribute");
    int per_cu = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tree,
                                                           static_cast<int>(kTreeBlock), lds),
              "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    return static_cast<uint32_t>(std::max(per_cu, 1) * std::max(ncu, 1));
}

void tree_launch(const TreeArgs& a, uint32_t nblocks, size_t lds, hipStream_t s) {
    LOKI_REQUIRE(a.items != nullptr && a.sync != nullptr && a.nitems >= 1,
                 "persistent batch: empty item list");
    LOKI_REQUIRE(nblocks >= 1 && lds <= 160 * 1024, "persistent batch: bad launch shape");
    k_tree<<<nblocks, kTreeBlock, lds, s>>>(a);
}

}  // namespace loki_amd::kern
