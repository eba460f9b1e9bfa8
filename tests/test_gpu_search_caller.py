"""The reference's only Python caller of the FFA path, ``loki.search.ffa_search``
(src/loki/search.py:7,41-57), with its import switched to ``loki_amd.libloki.ffa``: the same call
``compute_ffa_scores(ts_e, ts_v, cfg, quiet=, show_progress=)``, the same reshape
``(*plan.fold_shapes[-1][1:-2], len(cfg.score_widths))`` and the same parameter dictionary
``{"width": cfg.score_widths, **plan.params_dict}``, checked against the oracle."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TSAMP = 64e-6


def _search(ts_e, ts_v, cfg):
    """search.py:41-57 restated against the drop-in module (the Periodogram class comes from
    pyloki, which is not part of this path; its inputs are returned instead)."""
    from loki_amd.libloki.ffa import compute_ffa_scores

    snrs_flat, ffa_plan = compute_ffa_scores(ts_e, ts_v, cfg, quiet=True, show_progress=False)
    snrs = snrs_flat.reshape(*ffa_plan.fold_shapes[-1][1:-2], len(cfg.score_widths))
    params = {"width": cfg.score_widths, **ffa_plan.params_dict}
    return ffa_plan, params, snrs


@pytest.mark.parametrize(("lims", "nbins", "use_fourier"), [
    ([(-60000.0, 60000.0), (100.0, 110.0)], 32, False),
    ([(2.0, 8.0)], 64, False),
    ([(-2e5, 2e5), (-6e4, 6e4), (50.0, 55.0)], 32, False),
    ([(-60000.0, 60000.0), (100.0, 110.0)], 32, True),   # the reference's default domain
])
def test_ffa_search_call_shape(oracle, lims, nbins, use_fourier):
    from loki_amd.libloki.configs import PulsarSearchConfig
    from loki_amd.synth import make_series
    from oracle.plan import SearchConfig

    nsamps = 2**17
    cfg = PulsarSearchConfig(nsamps=nsamps, tsamp=TSAMP, nbins=nbins, eta=1.0,
                             param_limits=np.array(lims), use_fourier=use_fourier)
    ocfg = SearchConfig(nsamps, TSAMP, nbins, 1.0, lims, use_fourier=use_fourier)
    f0 = lims[-1][0] + 0.37 * (lims[-1][1] - lims[-1][0])
    ts_e, ts_v, _ = make_series(nsamps, TSAMP, 5, f0=f0, fwhm=0.05, snr=20.0, nbins=nbins)
    plan, params, snrs = _search(ts_e, ts_v, cfg)
    want, oplan = oracle.compute_ffa_scores(ts_e, ts_v, ocfg)
    np.testing.assert_array_equal(np.asarray(cfg.score_widths), ocfg.score_widths)
    assert snrs.shape == (*oplan.fold_shapes[-1][1:-2], len(ocfg.score_widths))
    if use_fourier:
        np.testing.assert_allclose(snrs.reshape(-1), want, rtol=1e-5, atol=1e-4)
    else:
        np.testing.assert_array_equal(snrs.reshape(-1), want)
    assert list(params) == ["width", *ocfg.param_names]
    for name, grid in oplan.params_dict.items():
        np.testing.assert_array_equal(params[name], grid)
    # the periodogram's lookup: the best (params..., width) cell finds the injected frequency
    best = np.unravel_index(int(np.argmax(snrs)), snrs.shape)
    f_best = params["freq"][best[-2]]
    df = params["freq"][1] - params["freq"][0]
    assert abs(f_best - f0) <= 3 * df


def test_ffa_classes_execute_like_the_cpu_module(oracle):
    """libloki.ffa.FFATime(cfg).execute(ts_e, ts_v, fold) fills a buffer_size fold whose first
    fold_size elements are the final level (loki_templates.hpp:47-61, ffa.cpp:83-85)."""
    from loki_amd.libloki import ffa
    from loki_amd.libloki.configs import PulsarSearchConfig
    from oracle.plan import SearchConfig

    lims = [(-60000.0, 60000.0), (100.0, 110.0)]
    cfg = PulsarSearchConfig(nsamps=2**16, tsamp=TSAMP, nbins=32, eta=1.0,
                             param_limits=np.array(lims), use_fourier=False)
    ocfg = SearchConfig(2**16, TSAMP, 32, 1.0, lims, use_fourier=False)
    rng = np.random.default_rng(3)
    ts_e = rng.standard_normal(2**16, dtype=np.float32)
    ts_v = np.ones(2**16, np.float32)
    obj = ffa.FFATime(cfg, show_progress=False)
    fold = np.zeros(obj.plan.buffer_size, np.float32)
    obj.execute(ts_e, ts_v, fold)
    want, _ = oracle.compute_ffa(ts_e, ts_v, ocfg)
    np.testing.assert_array_equal(fold[: obj.plan.fold_size], want)
    got, plan = ffa.compute_ffa_time(ts_e, ts_v, cfg, quiet=True)
    np.testing.assert_array_equal(got, want)
    with pytest.raises(RuntimeError):
        obj.execute(ts_e, ts_v, np.zeros(10, np.float32))
