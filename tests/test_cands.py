"""Candidate tables (loki_amd/cands.py): index -> parameter rows and the result archive (CPU)."""

from __future__ import annotations

import numpy as np
import pytest


def _plan():
    from loki_amd import libloki

    cfg = libloki.configs.PulsarSearchConfig(
        nsamps=2**17, tsamp=64e-6, nbins=32, eta=1.0,
        param_limits=np.array([(-60000.0, 60000.0), (100.0, 110.0)]), use_fourier=False)
    return cfg, libloki.plans.FFAPlanTime(cfg)


def test_param_sets_follow_final_grid_cell_centres():
    from loki_amd import cands

    cfg, plan = _plan()
    widths = np.array([1, 2, 4, 8], np.uint64)
    na, nf = (int(c) for c in plan.param_counts[-1])
    # trial (ia, jf) with width index iw -> index (ia * nf + jf) * nw + iw
    picks = [(0, 0, 0), (3, 17, 2), (na - 1, nf - 1, 3)]
    idx = np.array([(ia * nf + jf) * 4 + iw for ia, jf, iw in picks], np.uint32)
    rows = cands.param_sets(plan, idx, widths)
    lo_a, hi_a = -60000.0, 60000.0
    lo_f, hi_f = 100.0, 110.0
    for (ia, jf, iw), row in zip(picks, rows):
        # psr_utils.hpp:283-293 cell centres, in the reference's operation order
        assert row[0] == lo_a + (hi_a - lo_a) / na * (ia + 0.5)
        assert row[1] == lo_f + (hi_f - lo_f) / nf * (jf + 0.5)
        assert row[2] == widths[iw]
        # and within half a cell of the plan's grid value
        ga, gf = plan.params_dict.values()
        assert abs(row[0] - ga[ia]) < 1e-9 * (hi_a - lo_a) and abs(row[1] - gf[jf]) < 1e-9
    with pytest.raises(ValueError):
        cands.param_sets(plan, np.array([na * nf * 4], np.uint32), widths)


def test_result_archive_round_trip(tmp_path):
    from loki_amd import cands

    cfg, plan = _plan()
    widths = np.array([1, 2, 4], np.uint64)
    idx = np.array([5, 11, 40], np.uint32)
    snr = np.array([6.5, 7.25, 9.0], np.float32)
    ps = cands.param_sets(plan, idx, widths)
    path = cands.write_results(tmp_path / "run_ffa_results", list(cfg.param_names), widths, snr,
                               ps)
    got = cands.read_results(path)
    assert str(got["ffa_version"]) == cands.FFA_VERSION
    assert list(got["param_names"]) == ["accel", "freq", "width"]
    np.testing.assert_array_equal(got["scoring_widths"], widths)
    np.testing.assert_array_equal(got["snr"], snr)
    np.testing.assert_array_equal(got["param_sets"], ps)
    with pytest.raises(ValueError):
        cands.write_results(tmp_path / "bad", ["freq"], widths, snr, ps)
