"""Oracle restatement of the frequency-sweep region planner.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows lib/regions.cpp (pravirkr/loki v0.0.3): octave bands of the period range with a fixed
physical bin width (generate_ffa_regions, regions.cpp:31-77), the memory model of a sweep's
workspace (FFARegionStats, regions.cpp:81-152, byte counts turned into GiB in binary32 as
there) and the drift-aware, memory-bounded chunking of each band (FFARegionPlanner,
regions.cpp:189-414).  The plan sizes come from oracle/plan.py's Plan.

Pinned by tests/test_regions.py: closed-form bands (nbins = nbins_0 * 2^k, eta = eta_0 * 2^k
when p_min / tsamp bounds nbins), hand-computed memory figures, and the planner's own
invariants (contiguous cover, every chunk fits, no chunk can grow by one tolerance and still
fit, drift-widened limits).
"""

from __future__ import annotations

import dataclasses

import numpy as np

from .plan import KCVAL, Plan, SearchConfig

WRITE_BATCH = 1 << 16  # kFFAFreqSweepWriteBatchSize (ffa_freq_sweep.hpp)
_GIB = np.float32(1 << 30)


@dataclasses.dataclass
class Region:  # coord.hpp:28-33
    f_start: float
    f_end: float
    nbins: int
    eta: float


def generate_ffa_regions(p_min, p_max, tsamp, nbins_min, eta_min, octave_scale=2.0,
                         nbins_max=1024):
    """regions.cpp:31-77."""
    if not (p_min > 0 and p_max > p_min and tsamp > 0 and nbins_min >= 2 and eta_min > 0
            and octave_scale >= 1.0 and nbins_max >= nbins_min):
        raise ValueError("invalid region arguments")
    nb = min(int(nbins_min), int(p_min / tsamp))
    rho = eta_min / float(nb)
    bin_width = p_min / float(nb)
    out = []
    p_lo = p_min
    while p_lo < p_max:
        if nb >= nbins_max:
            out.append(Region(1.0 / p_max, 1.0 / p_lo, int(nbins_max),
                              float(np.round(rho * float(nbins_max)))))
            break
        p_hi = min(p_lo * octave_scale, p_max)
        out.append(Region(1.0 / p_hi, 1.0 / p_lo, nb, rho * float(nb)))
        p_lo = p_hi
        nb = int(p_lo / bin_width)
    return out


@dataclasses.dataclass
class RegionStats:  # regions.cpp:81-152
    max_buffer_size: int
    max_coord_size: int
    max_ncoords: int
    max_ffa_levels: int
    n_widths: int
    n_params: int
    n_samps: int
    max_passing_candidates: int
    use_fourier: bool
    use_gpu: bool = False

    @property
    def max_buffer_size_time(self):
        return 2 * self.max_buffer_size if self.use_fourier else self.max_buffer_size

    @property
    def max_scores_size(self):
        return max(self.max_ncoords * self.n_widths, self.max_passing_candidates)

    @property
    def write_param_sets_size(self):
        return WRITE_BATCH * (self.n_params + 1)

    @property
    def buffer_memory_usage(self):
        elem = 8 if self.use_fourier else 4
        return np.float32(2 * self.max_buffer_size * elem) / _GIB

    @property
    def coord_memory_usage(self):
        elem = 8 if self.n_params == 1 else 16
        return np.float32(self.max_coord_size * elem) / _GIB

    @property
    def extra_memory_usage(self):
        return np.float32(self.write_param_sets_size * 8 + self.max_scores_size * 8) / _GIB

    @property
    def cpu_memory_usage(self):
        return self.buffer_memory_usage + self.coord_memory_usage + self.extra_memory_usage

    @property
    def device_memory_usage(self):
        extra = np.float32((2 * self.n_samps + self.max_scores_size) * 4
                           + self.max_scores_size * 4) / _GIB
        return extra + self.buffer_memory_usage + self.coord_memory_usage

    @property
    def freq_sweep_memory_usage(self):
        return self.device_memory_usage if self.use_gpu else self.cpu_memory_usage


def max_drift(nparams, param_limits, tobs):
    """regions.cpp:189-213: the largest fractional Doppler shift over half the observation."""
    if nparams <= 1:
        return 0.0
    amax = [max(abs(lo), abs(hi)) for lo, hi in param_limits]
    half = tobs / 2.0
    if nparams == 2:
        return amax[0] * half / KCVAL
    if nparams == 3:
        return (amax[1] * half + amax[0] * half * half / 2.0) / KCVAL
    raise ValueError("Unsupported number of parameters for drift calculation")


def plan_regions(base: dict, n_widths: int, fourier: bool, max_process_memory_gb: float,
                 octave_scale: float = 2.0, max_passing_candidates: int = 1 << 20,
                 use_gpu: bool = True):
    """regions.cpp:215-414.  `base` holds the oracle SearchConfig's arguments; returns
    (chunk configs, chunk stats dicts, RegionStats of the maxima)."""
    cfg0 = SearchConfig(**base)
    regions = generate_ffa_regions(1.0 / cfg0.f_max, 1.0 / cfg0.f_min, cfg0.tsamp, cfg0.nbins,
                                   cfg0.eta, octave_scale, cfg0.nbins_max)
    drift = max_drift(cfg0.nparams, cfg0.param_limits, cfg0.tobs)
    if not 0.0 <= drift < 1.0:
        raise ValueError("max_drift must be in [0, 1)")
    limit = max_process_memory_gb - 0.5
    if limit <= 0.0:
        raise ValueError("max_process_memory_gb must exceed the safety margin")
    mx = [0, 0, 0, 0]
    cfgs, chunks = [], []

    def stats_of(b, c, n, lv):
        # the buffer element follows the config's use_fourier, the plan the fold kind
        return RegionStats(b, c, n, lv, n_widths, cfg0.nparams, cfg0.nsamps,
                           max_passing_candidates, cfg0.use_fourier, use_gpu)

    for r in regions:
        if r.f_end <= r.f_start:
            continue
        if r.f_start * (1.0 - drift) <= 0.0:
            raise ValueError("drift expansion would push f_start to a non-positive frequency")
        tol = max(1.0e-2, 1.0e-4 * (r.f_end - r.f_start))

        def evaluate(lo, hi, r=r):
            a, b = lo * (1.0 - drift), hi * (1.0 + drift)
            lim = list(cfg0.param_limits)
            lim[-1] = (a, b)
            args = dict(base, nbins=r.nbins, eta=r.eta, param_limits=lim)
            cfg = SearchConfig(**args)
            plan = Plan(cfg, fourier)
            sz = (plan.buffer_size, plan.coord_size, plan.ncoords[-1], plan.n_levels)
            own = stats_of(*sz).freq_sweep_memory_usage
            alloc = stats_of(*(max(m, s) for m, s in zip(mx, sz))).freq_sweep_memory_usage
            return cfg, sz, float(own), float(alloc)

        def fits(e):
            return e[3] <= limit

        hi = r.f_end
        while hi > r.f_start:
            best = evaluate(r.f_start, hi)
            lo = r.f_start
            if not fits(best):
                min_w = min(hi - r.f_start, 1.0e-2)
                best = evaluate(hi - min_w, hi)
                if not fits(best):
                    raise ValueError("cannot fit the minimum viable chunk")
                w_ok, w_bad = min_w, hi - r.f_start
                step = 0
                while step < 50 and (w_bad - w_ok) > tol:
                    w = (w_ok + w_bad) / 2.0  # std::midpoint of two finite doubles
                    e = evaluate(hi - w, hi)
                    if fits(e):
                        w_ok, best = w, e
                    else:
                        w_bad = w
                    step += 1
                lo = hi - w_ok
            rest = lo - r.f_start
            if 0.0 < rest <= 4.0 * tol:
                merged = evaluate(r.f_start, hi)
                if fits(merged):
                    lo, best = r.f_start, merged
            a, b = lo * (1.0 - drift), hi * (1.0 + drift)
            cfgs.append(best[0])
            chunks.append(dict(nominal_f_start=lo, nominal_f_end=hi, actual_f_start=a,
                               actual_f_end=b, nominal_width=hi - lo, actual_width=b - a,
                               total_memory_gb=best[2],
                               overlap_fraction=((b - a) - (hi - lo)) / (b - a)))
            mx = [max(m, s) for m, s in zip(mx, best[1])]
            hi = lo
    return cfgs, chunks, stats_of(*mx)
