"""Committed fixtures (tests/golden/ffa_small.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces every fixture bit for bit (a regression lock on the checker).
GPU: the HIP path reproduces the time-domain fixtures bit for bit and the Fourier fixtures to the
bound of tests/test_gpu_fourier.py, without running the oracle.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "ffa_small.npz"
NAMES = ["freq", "accel", "jerk"]
TSAMP = 64e-6


def _load(name):
    d = np.load(GOLDEN, allow_pickle=False)
    return {k[len(name) + 1:]: d[k] for k in d.files if k.startswith(name + "_")}


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(oracle, name):
    from oracle.plan import SearchConfig

    g = _load(name)
    lims = [tuple(r) for r in g["param_limits"]]
    cfg_t = SearchConfig(g["ts_e"].size, TSAMP, int(g["nbins"]), 1.0, lims, use_fourier=False)
    fold_t, _ = oracle.compute_ffa(g["ts_e"], g["ts_v"], cfg_t, nthreads=1)
    np.testing.assert_array_equal(fold_t, g["fold_time"])
    scores, _ = oracle.compute_ffa_scores(g["ts_e"], g["ts_v"], cfg_t, widths=g["widths"],
                                          nthreads=1)
    np.testing.assert_array_equal(scores, g["scores_time"])
    cfg_f = SearchConfig(g["ts_e"].size, TSAMP, int(g["nbins"]), 1.0, lims, use_fourier=True)
    fold_f, _ = oracle.compute_ffa(g["ts_e"], g["ts_v"], cfg_f, fourier=True, nthreads=1)
    np.testing.assert_array_equal(fold_f, g["fold_fourier"])


def _cfg(g, fourier):
    from loki_amd import libloki

    return libloki.configs.PulsarSearchConfig(
        nsamps=g["ts_e"].size, tsamp=TSAMP, nbins=int(g["nbins"]), eta=1.0,
        param_limits=g["param_limits"], use_fourier=fourier)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(name):
    from loki_amd import libculoki

    g = _load(name)
    fold, _ = libculoki.ffa.compute_ffa_time_cuda(g["ts_e"], g["ts_v"], _cfg(g, False))
    np.testing.assert_array_equal(fold, g["fold_time"])
    scores, _ = libculoki.ffa.compute_ffa_scores_cuda(g["ts_e"], g["ts_v"], _cfg(g, False),
                                                      widths=g["widths"])
    np.testing.assert_array_equal(scores, g["scores_time"])
    fold_f, _ = libculoki.ffa.compute_ffa_fourier_cuda(g["ts_e"], g["ts_v"], _cfg(g, True))
    nbf = int(g["nbins"]) // 2 + 1
    got = fold_f.reshape(-1, nbf).astype(np.complex128)
    want = g["fold_fourier"].reshape(-1, nbf).astype(np.complex128)
    err = np.abs(got - want).max(axis=1)
    assert np.all(err <= 1e-5 * np.linalg.norm(want, axis=1))
