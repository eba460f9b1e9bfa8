"""Region planner of the frequency sweep (lib/regions.cpp) on the host.

The reference ships no test for the planner, so the oracle (oracle/regions.py) is pinned by
closed forms and invariants, and the product is held equal to the oracle:

* bands: with p_min / tsamp < nbins_min the first band has nbins_0 = floor(p_min / tsamp) and
  every octave doubles nbins and eta exactly; bands tile [f_min, f_max] with ratio
  octave_scale; a cap at nbins_max makes the last band reach f_min with eta rounded;
* memory figures: hand-computed byte counts of a workspace;
* chunks: each band is covered contiguously from the top; every chunk fits the limit; no chunk
  that stops short of its band's bottom can grow by one boundary tolerance and still fit; the
  chunk configs search [lo (1 - drift), hi (1 + drift)] with drift the Doppler fraction of the
  accel/jerk ranges.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import regions as oreg
from oracle.plan import KCVAL, Plan, SearchConfig


def _lib():
    from loki_amd import libloki

    return libloki


def test_regions_closed_form_doubling():
    p_min, p_max, tsamp = 1 / 500.0, 1.0, 64e-6
    got = _lib().plans.generate_ffa_regions(p_min, p_max, tsamp, 64, 1.0, nbins_max=10**6)
    n0 = int(p_min / tsamp)  # 31 < 64
    assert n0 == 31
    k = np.arange(got.size)
    np.testing.assert_array_equal(got["nbins"], n0 * 2**k)
    np.testing.assert_array_equal(got["eta"], 2.0**k)
    np.testing.assert_allclose(got["f_end"], 500.0 / 2.0**k, rtol=1e-15)
    assert got["f_start"][-1] == 1.0 and got["f_end"][0] == 500.0
    np.testing.assert_array_equal(got["f_start"][:-1], got["f_end"][1:])
    np.testing.assert_allclose(got["f_end"][:-1] / got["f_start"][:-1], 2.0, rtol=1e-15)
    assert got.size == math.ceil(math.log2(500.0))


@pytest.mark.parametrize("args", [
    (1 / 500.0, 1.0, 64e-6, 64, 1.0, 2.0, 1024),
    (1 / 200.0, 2.0, 1e-4, 32, 2.0, 2.0, 256),
    (1 / 97.0, 3.3, 5e-5, 50, 1.5, 1.7, 400),
    (1 / 1000.0, 1.0, 1e-6, 128, 1.0, 3.0, 1024),
    (0.01, 0.012, 1e-4, 16, 1.0, 2.0, 16),
])
def test_regions_match_oracle(args):
    got = _lib().plans.generate_ffa_regions(*args[:5], growth_factor=args[5], nbins_max=args[6])
    want = oreg.generate_ffa_regions(*args)
    assert got.size == len(want)
    for g, w in zip(got, want):
        assert (g["f_start"], g["f_end"], g["nbins"], g["eta"]) == (
            w.f_start, w.f_end, w.nbins, w.eta)
    # the cap: the last band runs to f_min with nbins_max bins
    if want[-1].nbins == args[6]:
        assert want[-1].f_start == 1.0 / args[1]
        assert want[-1].eta == round(args[4] / min(args[3], int(args[0] / args[2])) * args[6])


def test_regions_errors():
    gen = _lib().plans.generate_ffa_regions
    with pytest.raises(RuntimeError, match="p_max"):
        gen(1.0, 0.5, 1e-4, 64, 1.0)
    with pytest.raises(RuntimeError, match="octave_scale"):
        gen(0.01, 1.0, 1e-4, 64, 1.0, growth_factor=0.5)
    with pytest.raises(RuntimeError, match="nbins_max"):
        gen(0.01, 1.0, 1e-4, 64, 1.0, nbins_max=32)


def test_region_stats_hand_computed():
    st = _lib().plans.FFARegionStats(max_buffer_size=3 << 28, max_coord_size=1000,
                                     max_ncoords=5000, max_ffa_levels=9, n_widths=7, n_params=2,
                                     n_samps=1 << 24, max_passing_candidates=1 << 20,
                                     use_fourier=True, use_gpu=True)
    assert st.max_buffer_size_time == 6 << 28
    assert st.max_scores_size == 1 << 20  # max(5000 * 7, 2^20)
    assert st.write_param_sets_size == (1 << 16) * 3
    gib = float(1 << 30)
    assert st.buffer_memory_usage == 2 * (3 << 28) * 8 / gib == 12.0
    assert st.coord_memory_usage == np.float32(np.float32(16000) / np.float32(gib))
    extra = np.float32(np.float32((1 << 16) * 3 * 8 + (1 << 20) * 8) / np.float32(gib))
    assert st.extra_memory_usage == extra
    dev = np.float32(((2 << 24) + (1 << 20)) * 4 + (1 << 20) * 4) / np.float32(gib)
    assert st.freq_sweep_memory_usage == np.float32(dev + np.float32(12.0)
                                                    + st.coord_memory_usage)
    cpu = _lib().plans.FFARegionStats(3 << 28, 1000, 5000, 9, 7, 2, 1 << 24, 1 << 20, True)
    assert cpu.freq_sweep_memory_usage == np.float32(np.float32(12.0) + cpu.coord_memory_usage
                                                     + cpu.extra_memory_usage)
    o = oreg.RegionStats(3 << 28, 1000, 5000, 9, 7, 2, 1 << 24, 1 << 20, True, True)
    assert o.freq_sweep_memory_usage == st.freq_sweep_memory_usage


_CASES = {
    # name: (config args, memory limit GB)
    "freq_fits": (dict(nsamps=2**18, tsamp=1e-3, nbins=32, eta=1.0,
                       param_limits=[[1.0, 50.0]]), 8.0),
    "freq_split": (dict(nsamps=2**20, tsamp=1e-4, nbins=64, eta=1.0,
                        param_limits=[[2.0, 100.0]]), 0.56),
    "accel_split": (dict(nsamps=2**19, tsamp=1e-4, nbins=64, eta=1.0,
                         param_limits=[[-200.0, 200.0], [5.0, 80.0]]), 0.57),
    "jerk_split": (dict(nsamps=2**18, tsamp=2e-4, nbins=32, eta=1.0,
                        param_limits=[[-50.0, 50.0], [-100.0, 100.0], [4.0, 40.0]]), 0.51),
    "fourier_accel": (dict(nsamps=2**19, tsamp=1e-4, nbins=64, eta=1.0, use_fourier=True,
                           param_limits=[[-100.0, 100.0], [8.0, 60.0]]), 0.6),
}


def _configs(name):
    base, mem = _CASES[name]
    base = dict(base)
    fourier = bool(base.pop("use_fourier", False))
    lims = np.array(base["param_limits"], np.float64)
    cfg = _lib().configs.PulsarSearchConfig(
        nsamps=base["nsamps"], tsamp=base["tsamp"], nbins=base["nbins"], eta=base["eta"],
        param_limits=lims, use_fourier=fourier, max_process_memory_gb=mem,
        max_passing_candidates=1 << 14)
    obase = dict(base, use_fourier=fourier)
    return cfg, obase, fourier, mem


@pytest.mark.parametrize("name", list(_CASES))
def test_planner_matches_oracle(name):
    cfg, obase, fourier, mem = _configs(name)
    planner = (_lib().plans.FFARegionPlannerFourier if fourier
               else _lib().plans.FFARegionPlannerTime)(cfg)
    ocfgs, ochunks, ostats = oreg.plan_regions(obase, len(cfg.score_widths), fourier, mem,
                                               max_passing_candidates=1 << 14)
    assert planner.nregions == len(ocfgs)
    for c, oc, ch, och in zip(planner.cfgs, ocfgs, planner.chunk_stats, ochunks):
        assert c.nbins == oc.nbins and c.eta == oc.eta
        np.testing.assert_array_equal(c.param_limits, np.array(oc.param_limits))
        assert c.bseg_brute == oc.bseg_brute
        for k, v in och.items():
            assert ch[k] == (np.float32(v) if k == "total_memory_gb" else v), k
    st = planner.stats
    assert (st.max_buffer_size, st.max_coord_size, st.max_ncoords, st.max_ffa_levels) == (
        ostats.max_buffer_size, ostats.max_coord_size, ostats.max_ncoords,
        ostats.max_ffa_levels)
    assert st.freq_sweep_memory_usage == ostats.freq_sweep_memory_usage


@pytest.mark.parametrize("name", list(_CASES))
def test_planner_invariants(name):
    """The oracle's chunks obey the planner's contract (checked without the product)."""
    _, obase, fourier, mem = _configs(name)
    cfg0 = SearchConfig(**obase)
    nw = len(cfg0.score_widths)
    cfgs, chunks, stats = oreg.plan_regions(obase, nw, fourier, mem,
                                            max_passing_candidates=1 << 14)
    regions = oreg.generate_ffa_regions(1 / cfg0.f_max, 1 / cfg0.f_min, cfg0.tsamp, cfg0.nbins,
                                        cfg0.eta, 2.0, cfg0.nbins_max)
    amax = [max(abs(a), abs(b)) for a, b in cfg0.param_limits]
    half = cfg0.tobs / 2
    if cfg0.nparams == 1:
        drift = 0.0
    elif cfg0.nparams == 2:  # accel: v / c = a t / c
        drift = amax[0] * half / KCVAL
    else:  # jerk: a t + j t^2 / 2
        drift = (amax[1] * half + amax[0] * half**2 / 2) / KCVAL
    limit = mem - 0.5
    i = 0
    for r in regions:
        hi = r.f_end
        tol = max(1e-2, 1e-4 * (r.f_end - r.f_start))
        while hi > r.f_start:
            ch, cfg = chunks[i], cfgs[i]
            assert ch["nominal_f_end"] == hi and ch["nominal_f_start"] < hi
            assert cfg.nbins == r.nbins and cfg.eta == r.eta
            assert cfg.param_limits[-1] == (ch["nominal_f_start"] * (1 - drift), hi * (1 + drift))
            assert cfg.param_limits[:-1] == cfg0.param_limits[:-1]
            plan = Plan(cfg, fourier)
            own = oreg.RegionStats(plan.buffer_size, plan.coord_size, plan.ncoords[-1],
                                   plan.n_levels, nw, cfg0.nparams, cfg0.nsamps, 1 << 14,
                                   cfg0.use_fourier, True).freq_sweep_memory_usage
            assert own <= limit and ch["total_memory_gb"] == float(own)
            lo = ch["nominal_f_start"]
            if lo > r.f_start:
                # maximal: one more tolerance of width no longer fits on its own
                lo2 = max(r.f_start, lo - 2 * tol)
                lim = list(cfg.param_limits)
                lim[-1] = (lo2 * (1 - drift), hi * (1 + drift))
                p2 = Plan(SearchConfig(**dict(obase, nbins=r.nbins, eta=r.eta, param_limits=lim)),
                          fourier)
                st2 = oreg.RegionStats(max(stats.max_buffer_size, p2.buffer_size),
                                       max(stats.max_coord_size, p2.coord_size),
                                       max(stats.max_ncoords, p2.ncoords[-1]),
                                       stats.max_ffa_levels, nw, cfg0.nparams, cfg0.nsamps,
                                       1 << 14, cfg0.use_fourier, True)
                assert st2.freq_sweep_memory_usage > limit
            hi = lo
            i += 1
    assert i == len(chunks)
    if name.endswith("split"):
        assert len(chunks) > len(regions)


def test_planner_errors():
    lib = _lib()
    cfg = lib.configs.PulsarSearchConfig(nsamps=2**20, tsamp=1e-4, nbins=64, eta=1.0,
                                         param_limits=np.array([[2.0, 100.0]]),
                                         max_process_memory_gb=0.4)
    with pytest.raises(RuntimeError, match="margin"):
        lib.plans.FFARegionPlannerTime(cfg)
    cfg = lib.configs.PulsarSearchConfig(nsamps=2**20, tsamp=1e-4, nbins=64, eta=1.0,
                                         param_limits=np.array([[-1e9, 1e9], [2.0, 100.0]]))
    with pytest.raises(RuntimeError, match="max_drift"):
        lib.plans.FFARegionPlannerTime(cfg)
