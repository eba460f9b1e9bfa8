"""Pinning the FP64 refinement oracle (oracle/refine.py) on the CPU.

  * with fdot = fddot = 0 the continuous fold equals the oracle's Fourier brute fold over one
    segment spanning the series (itself pinned against a direct DFT and the time fold);
  * a pure tone with the chirp of theta folds to E_1 = N/2 only when the quadratic and cubic
    phase terms carry the factors 1/2 and 1/6 and the right signs;
  * the score equals the oracle's FP64 boxcar S/N of the inverse-transformed profile;
  * the analytic adjoint gradient equals central finite differences of the score.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import refine


def _series(n, seed, theta=None, tsamp=64e-6, t_ref=None, amp=0.0):
    rng = np.random.default_rng(seed)
    ts_e = rng.standard_normal(n)
    t_ref = n * tsamp / 2 if t_ref is None else t_ref
    if theta is not None and amp:
        _, phi = refine.phases(n, tsamp, t_ref, theta)
        fr = phi - np.floor(phi)
        d = np.minimum(fr, 1 - fr)
        ts_e += amp * np.exp(-0.5 * (d / 0.02) ** 2)
    return ts_e.astype(np.float32), np.ones(n, np.float32), tsamp, t_ref


def test_fold_without_derivatives_is_one_segment_brute_fold(oracle):
    ts_e, ts_v, tsamp, t_ref = _series(4096, 1)
    ts_v = (0.5 + np.random.default_rng(2).random(4096)).astype(np.float32)
    for f in (13.25, 97.0):
        E, V = refine.fold(ts_e, ts_v, tsamp, t_ref, (f, 0.0, 0.0), 32)
        bf = oracle.brute_fold_fourier(ts_e, ts_v, np.array([f]), 4096, 32, tsamp, t_ref)
        bf = bf.reshape(2, 17)
        np.testing.assert_allclose(E, bf[0], rtol=0, atol=2e-5 * np.abs(ts_e).sum())
        np.testing.assert_allclose(V, bf[1], rtol=0, atol=2e-5 * np.abs(ts_v).sum())


@pytest.mark.parametrize("theta", [(40.0, 4.0, 0.0), (40.0, 0.0, 10.0), (55.5, -3.0, 8.0)])
def test_chirped_tone_folds_coherently(theta):
    n, tsamp = 1 << 15, 64e-6
    t_ref = n * tsamp / 2
    _, phi = refine.phases(n, tsamp, t_ref, theta)
    # independent restatement of the chirp: instantaneous frequency integrated by trapezoids
    tau = np.arange(n) * tsamp - t_ref
    f_inst = theta[0] + theta[1] * tau + theta[2] * tau * tau / 2.0
    integ = np.concatenate([[0.0], np.cumsum((f_inst[1:] + f_inst[:-1]) * tsamp / 2.0)])
    k0 = int(np.argmin(np.abs(tau)))
    integ -= integ[k0] - phi[k0]
    assert np.max(np.abs(integ - phi)) < 1e-6
    tone = np.cos(2 * np.pi * phi).astype(np.float32)
    E, _ = refine.fold(tone, np.ones(n, np.float32), tsamp, t_ref, theta, 16)
    assert abs(E[1] - n / 2) < 1e-2 * n
    assert np.all(np.abs(E[2:]) < 1e-2 * n)
    # the plain (derivative-free) fold loses the chirped tone
    E0, _ = refine.fold(tone, np.ones(n, np.float32), tsamp, t_ref, (theta[0], 0.0, 0.0), 16)
    assert abs(E0[1]) < 0.5 * n


@pytest.mark.parametrize("width", [1, 3, 8])
def test_score_is_boxcar_snr_of_profile(oracle, width):
    theta = (37.3, 2e-4, -1e-5)
    ts_e, ts_v, tsamp, t_ref = _series(8192, 3, theta, amp=0.3)
    E, V = refine.fold(ts_e, ts_v, tsamp, t_ref, theta, 32)
    e, v = np.fft.irfft(E, n=32), np.fft.irfft(V, n=32)
    want = oracle.snr_boxcar_3d_f64(np.stack([e, v])[None], np.array([width], np.uint64), 32)
    got = refine.score(ts_e, ts_v, tsamp, t_ref, theta, 32, width)
    assert abs(got - float(want.reshape(-1)[0])) < 1e-9 * max(1.0, abs(got))


@pytest.mark.parametrize("seed,nbins,width", [(4, 32, 4), (5, 16, 2), (6, 64, 9)])
def test_adjoint_gradient_matches_central_differences(seed, nbins, width):
    n = 1 << 14
    tsamp = 64e-6
    T = n * tsamp
    theta_true = (37.3, 0.5 / T**2, 2.0 / T**3)
    ts_e, ts_v, tsamp, t_ref = _series(n, seed, theta_true, amp=0.4)
    theta = np.array([37.3 + 0.3 / T, 0.9 / T**2, 1.0 / T**3])
    s, g = refine.grad(ts_e, ts_v, tsamp, t_ref, theta, nbins, width)
    assert s == pytest.approx(refine.score(ts_e, ts_v, tsamp, t_ref, theta, nbins, width),
                              rel=1e-12)
    gfd = refine.grad_fd(ts_e, ts_v, tsamp, t_ref, theta, nbins, width, rel=1e-6)
    # relative step 1e-6 of each parameter; FP64 central differences are good to ~1e-6 here
    np.testing.assert_allclose(g, gfd, rtol=1e-5, atol=0)


def test_adjoint_gradient_sign_points_uphill():
    n = 1 << 14
    tsamp = 64e-6
    T = n * tsamp
    theta_true = np.array([51.0, 0.0, 0.0])
    ts_e, ts_v, tsamp, t_ref = _series(n, 8, theta_true, amp=0.5)
    theta = theta_true + np.array([0.2 / T, 0.0, 0.0])
    s0, g = refine.grad(ts_e, ts_v, tsamp, t_ref, theta, 32, 2)
    assert g[0] < 0  # the peak lies at lower f
    step = 0.05 / T
    s1 = refine.score(ts_e, ts_v, tsamp, t_ref, theta - np.array([step, 0, 0]), 32, 2)
    assert s1 > s0
