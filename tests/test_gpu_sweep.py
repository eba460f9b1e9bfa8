"""Frequency sweep on the GPU (ffa_freq_sweep_cuda.cu:21-365 FFAFreqSweepCUDA).

The sweep runs the region planner's chunks (pinned on the host in tests/test_regions.py) one
after another, each as FFA + score-and-filter at cfg.snr_min with the base config's boxcar
widths, into one candidate list.  The candidates must equal, chunk by chunk, the thresholded
full scores of that chunk's config (which the parity tests hold to the oracle), with the rows
decoded to the chunk's final-grid cell centres (psr_utils.hpp:283-293); one time-domain chunk
is also checked against the oracle directly.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSAMPS, TSAMP = 2**17, 2e-4
SIGNALS = [
    dict(f0=13.37, fdot=0.0, fddot=0.0, fwhm=0.05, snr=25.0),
    dict(f0=27.1, fdot=6.0e-7, fddot=0.0, fwhm=0.08, snr=18.0),  # accel ~ 6.6 m/s^2
    dict(f0=6.3, fdot=0.0, fddot=0.0, fwhm=0.1, snr=20.0),
]


def _config(fourier, mem=0.503, snr_min=7.0):
    from loki_amd import libloki

    return libloki.configs.PulsarSearchConfig(
        nsamps=NSAMPS, tsamp=TSAMP, nbins=32, eta=1.0,
        param_limits=np.array([[-10.0, 10.0], [5.0, 40.0]]), use_fourier=fourier,
        max_process_memory_gb=mem, snr_min=snr_min, max_passing_candidates=1 << 12)


def _series():
    from loki_amd.synth import make_injected_series

    ts_e, ts_v, _ = make_injected_series(NSAMPS, TSAMP, 77, SIGNALS, nbins=32)
    return ts_e, ts_v


def _cell_rows(cfg, idx, widths):
    """Rows (params..., width) of the reference's save_results (ffa_freq_sweep_cuda.cu:300-322)."""
    from loki_amd import libloki
    from oracle.plan import param_val_at_idx

    plan = (libloki.plans.FFAPlanFourier if cfg.use_fourier else libloki.plans.FFAPlanTime)(cfg)
    counts = [int(c) for c in plan.param_counts[-1]]
    nw = widths.size
    cols = np.unravel_index(idx.astype(np.int64) // nw, counts)
    rows = [param_val_at_idx(lo, hi, c, col)
            for (lo, hi), c, col in zip(cfg.param_limits, counts, cols)]
    rows.append(widths[idx.astype(np.int64) % nw].astype(np.float64))
    return np.stack(rows, axis=1)


@pytest.mark.parametrize("fourier", [False, True])
def test_sweep_equals_per_chunk_filtered_scores(fourier, tmp_path):
    from loki_amd import cands, libculoki

    cfg = _config(fourier)
    ts_e, ts_v = _series()
    sweep = libculoki.ffa.FFAFreqSweepCUDA(cfg, device_id=0)
    planner = sweep.planner
    assert planner.nregions == 9  # three octaves, three chunks each at this memory limit
    path = sweep.execute(ts_e, ts_v, str(tmp_path), file_prefix="sweep")
    assert path.name == "sweep_ffa_results.npz"
    got = cands.read_results(path)
    assert list(got["param_names"]) == list(cfg.param_names) + ["width"]
    widths = np.asarray(cfg.score_widths, np.uint64)
    np.testing.assert_array_equal(got["scoring_widths"], widths)

    want_snr, want_rows = [], []
    cls = libculoki.ffa.FFACOMPLEXCUDA if fourier else libculoki.ffa.FFACUDA
    for k, ccfg in enumerate(planner.cfgs):
        full = cls(ccfg, device_id=0).scores(ts_e, ts_v, widths)
        idx = np.nonzero(full > np.float32(cfg.snr_min))[0].astype(np.uint32)
        want_snr.append(full[idx])
        want_rows.append(_cell_rows(ccfg, idx, widths))
        print(f"chunk {k}: nbins {ccfg.nbins} f {ccfg.param_limits[-1]} -> {idx.size}", flush=True)
    want_snr = np.concatenate(want_snr)
    want_rows = np.concatenate(want_rows)
    assert want_snr.size > 50  # the injected pulses pass
    np.testing.assert_array_equal(got["snr"], want_snr)
    np.testing.assert_array_equal(got["param_sets"], want_rows)
    snr, rows = sweep.results()
    np.testing.assert_array_equal(snr, want_snr)
    np.testing.assert_array_equal(rows, want_rows)
    # every injected pulse is found near its frequency with most of its S/N
    for sg in SIGNALS:
        near = np.abs(rows[:, -2] - sg["f0"]) < 0.05
        assert near.any() and snr[near].max() > 0.6 * sg["snr"], sg


def test_sweep_chunk_equals_oracle(tmp_path):
    from loki_amd import libculoki
    from oracle import oracle
    from oracle.plan import SearchConfig

    cfg = _config(False)
    ts_e, ts_v = _series()
    sweep = libculoki.ffa.FFAFreqSweepCUDA(cfg, device_id=0)
    sweep.execute(ts_e, ts_v, str(tmp_path), file_prefix="oracle_chunk")
    snr, rows = sweep.results()
    per = np.asarray(sweep.per_region, np.int64)
    assert per.sum() == snr.size
    widths = np.asarray(cfg.score_widths, np.uint64)
    ccfgs = sweep.planner.cfgs
    # the chunk holding the 13.37 Hz pulse
    k = next(i for i, c in enumerate(ccfgs)
             if c.param_limits[-1][0] <= 13.37 <= c.param_limits[-1][1])
    c = ccfgs[k]
    ocfg = SearchConfig(NSAMPS, TSAMP, int(c.nbins), float(c.eta), c.param_limits.tolist(),
                        use_fourier=False)
    full, _ = oracle.compute_ffa_scores(ts_e, ts_v, ocfg, widths)
    idx = np.nonzero(full > np.float32(cfg.snr_min))[0]
    lo, hi = per[:k].sum(), per[: k + 1].sum()
    assert hi - lo == idx.size > 10
    np.testing.assert_array_equal(snr[lo:hi], full[idx].astype(np.float32))
    np.testing.assert_array_equal(rows[lo:hi], _cell_rows(c, idx.astype(np.uint32), widths))


def test_sweep_rejects_wrong_length_and_full_buffer(tmp_path):
    from loki_amd import libculoki

    ts_e, ts_v = _series()
    sweep = libculoki.ffa.FFAFreqSweepCUDA(_config(False), device_id=0)
    with pytest.raises(RuntimeError, match="nsamps"):
        sweep.execute(ts_e[:-2], ts_v[:-2], str(tmp_path), "bad")
    # a threshold every trial passes overflows the candidate list at the second chunk
    low = libculoki.ffa.FFAFreqSweepCUDA(_config(False, snr_min=-1e9), device_id=0)
    with pytest.raises(RuntimeError, match="candidate buffer full"):
        low.execute(ts_e, ts_v, str(tmp_path), "full")
    assert not (tmp_path / "full_ffa_results.npz").exists()
