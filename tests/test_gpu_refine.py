"""Differentiable Fourier-domain candidate refinement on the GPU against the FP64 oracle.

Forward scores are checked against oracle/refine.py (FP64 fold, inverse DFT and boxcar) to the
Fourier-path S/N bound |dS| <= 1e-4 + 1e-5 |S|.  Gradients of the adjoint kernel are checked
against the oracle's FP64 adjoint (itself pinned against FP64 central differences at relative
step 1e-6 in tests/test_oracle_refine.py) and against central differences of the GPU forward.
Gradients are compared in turn units g_u = dS/dtheta / phase_scale (dS per turn of phase moved
at the series' far end), where the three parameters are commensurable.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TSAMP = 64e-6


def _torch():
    import torch

    return torch


def _inject(n, seed, thetas, amps, fwhm=0.04):
    """N(0,1) noise plus Gaussian pulses at the continuous ephemerides `thetas` (FP64)."""
    from oracle import refine as oref

    rng = np.random.default_rng(seed)
    ts = rng.standard_normal(n)
    t_ref = n * TSAMP / 2
    sig = fwhm / 2.3548200450309493
    for th, a in zip(thetas, amps):
        _, phi = oref.phases(n, TSAMP, t_ref, th)
        fr = phi - np.floor(phi)
        d = np.minimum(fr, 1.0 - fr)
        ts += a * np.exp(-0.5 * (d / sig) ** 2)
    return ts.astype(np.float32), np.ones(n, np.float32), t_ref


def _dev(*arrs):
    torch = _torch()
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") for a in arrs]


def _candidates(n, seed, ncand, theta_true):
    """Candidates scattered within a few bins of phase around theta_true and far from it."""
    from loki_amd import refine

    rng = np.random.default_rng(seed + 1000)
    scale = np.array(refine.phase_scale(n, TSAMP, n * TSAMP / 2))
    du = rng.uniform(-0.15, 0.15, (ncand, 3))
    th = np.asarray(theta_true)[None, :] + du / scale[None, :]
    th[-1, 0] += 7.3 / scale[0]  # off the signal: noise only
    return th, scale


CASES = [
    # (log2 nsamps, nbins, width, seed)
    (16, 32, 4, 11),
    (16, 64, 6, 12),
    (15, 16, 2, 13),
    (16, 128, 17, 14),
    (17, 64, 1, 15),
]


@pytest.mark.parametrize("lg,nbins,width,seed", CASES)
def test_scores_and_gradients_match_fp64_oracle(lg, nbins, width, seed):
    from loki_amd import refine
    from oracle import refine as oref

    torch = _torch()
    n = 1 << lg
    T = n * TSAMP
    theta_true = (61.7, 0.8 / T**2, -3.0 / T**3)
    ts_e, ts_v, t_ref = _inject(n, seed, [theta_true], [0.08])
    th, scale = _candidates(n, seed, 8, theta_true)
    de, dv, dth = _dev(ts_e, ts_v, th)
    s_gpu, state = refine.score_raw(de, dv, dth, TSAMP, t_ref, nbins, width)
    g_gpu = refine.grad_raw(de, dv, dth, state, TSAMP, t_ref, nbins)
    torch.cuda.synchronize()
    s_gpu, g_gpu = s_gpu.cpu().numpy().astype(np.float64), g_gpu.cpu().numpy()
    worst_s, worst_g = 0.0, 0.0
    for c in range(th.shape[0]):
        s_ref, g_ref = oref.grad(ts_e, ts_v, TSAMP, t_ref, th[c], nbins, width)
        ds = abs(s_gpu[c] - s_ref)
        assert ds <= 1e-4 + 1e-5 * abs(s_ref), (c, s_gpu[c], s_ref)
        gu_ref, gu = g_ref / scale, g_gpu[c] / scale
        err = np.linalg.norm(gu - gu_ref)
        # FP32 per-sample adjoint terms summed in FP64: 1e-4 of the gradient norm plus a floor
        # of 2e-3 S/N per turn for gradients that vanish at the peak
        assert err <= 1e-4 * np.linalg.norm(gu_ref) + 2e-3, (c, gu, gu_ref)
        if c < 2:
            # the oracle's central differences at relative step 1e-6 (FP64 forward); their own
            # O(h^2) truncation error reaches ~2e-4 of the gradient for width-1 peaks
            gfd = oref.grad_fd(ts_e, ts_v, TSAMP, t_ref, th[c], nbins, width, rel=1e-6)
            assert np.linalg.norm(gu - gfd / scale) <= 1e-3 * np.linalg.norm(gu_ref) + 2e-3
        worst_s = max(worst_s, ds)
        worst_g = max(worst_g, err / (np.linalg.norm(gu_ref) + 1e-30))
    print(f"n=2^{lg} nbins={nbins} w={width}: max|dS|={worst_s:.2e} max rel|dg|={worst_g:.2e}")


def test_ts_v_enters_the_score():
    """Non-uniform variance series: scores and gradients still follow the oracle."""
    from loki_amd import refine
    from oracle import refine as oref

    n, nbins, width = 1 << 15, 32, 3
    T = n * TSAMP
    theta_true = (23.4, 0.0, 0.0)
    ts_e, _, t_ref = _inject(n, 21, [theta_true], [0.1])
    ts_v = (0.5 + np.random.default_rng(22).random(n)).astype(np.float32)
    th = np.array([[23.4 + 0.05 / T, 0.1 / T**2, 0.2 / T**3], [23.4, 0.0, 0.0]])
    de, dv, dth = _dev(ts_e, ts_v, th)
    s, st = refine.score_raw(de, dv, dth, TSAMP, t_ref, nbins, width)
    g = refine.grad_raw(de, dv, dth, st, TSAMP, t_ref, nbins).cpu().numpy()
    s = s.cpu().numpy()
    scale = np.array(refine.phase_scale(n, TSAMP, t_ref))
    for c in range(2):
        s_ref, g_ref = oref.grad(ts_e, ts_v, TSAMP, t_ref, th[c], nbins, width)
        assert abs(s[c] - s_ref) <= 1e-4 + 1e-5 * abs(s_ref)
        assert np.linalg.norm((g[c] - g_ref) / scale) <= \
            1e-4 * np.linalg.norm(g_ref / scale) + 2e-3


def _gpu_fd(de, dv, th, t_ref, nbins, width, steps):
    """Central differences of the GPU forward, one parameter at a time, for every candidate."""
    from loki_amd import refine

    torch = _torch()
    out = np.zeros((th.shape[0], 3))
    for i in range(3):
        tp, tm = th.copy(), th.copy()
        tp[:, i] += steps[:, i]
        tm[:, i] -= steps[:, i]
        sp, _ = refine.score_raw(de, dv, torch.from_numpy(tp).cuda(), TSAMP, t_ref, nbins, width,
                                 want_state=False)
        sm, _ = refine.score_raw(de, dv, torch.from_numpy(tm).cuda(), TSAMP, t_ref, nbins, width,
                                 want_state=False)
        out[:, i] = (sp.double() - sm.double()).cpu().numpy() / (2.0 * steps[:, i])
    return out


def test_full_size_gradient_matches_gpu_central_differences():
    """2^22 samples, nbins 64 (33 harmonics), the Fourier-domain refinement workload: the
    adjoint against central differences of the FP32 forward with steps of 1e-3 turns at the
    far end.  (A relative step of 1e-6 moves f by 0.025 turns here, too far for a central
    difference of a peak 1/64 turn wide, and fdot, fddot by far less than the forward's FP32
    resolution; that step is applied to the FP64 oracle at smaller sizes.)"""
    from loki_amd import refine

    torch = _torch()
    n, nbins, width = 1 << 22, 64, 5
    T = n * TSAMP
    theta_true = (187.25, 4e-9, -5e-15)
    ts_e, ts_v, t_ref = _inject(n, 31, [theta_true], [0.03])
    scale = np.array(refine.phase_scale(n, TSAMP, t_ref))
    rng = np.random.default_rng(32)
    th = np.asarray(theta_true)[None, :] + rng.uniform(-0.05, 0.05, (6, 3)) / scale[None, :]
    de, dv, dth = _dev(ts_e, ts_v, th)
    s, st = refine.score_raw(de, dv, dth, TSAMP, t_ref, nbins, width)
    g = refine.grad_raw(de, dv, dth, st, TSAMP, t_ref, nbins).cpu().numpy()
    steps = np.empty_like(th)
    steps[:] = 1e-3 / scale[None, :]
    gfd = _gpu_fd(de, dv, th, t_ref, nbins, width, steps)
    torch.cuda.synchronize()
    for c in range(th.shape[0]):
        gu, gfu = g[c] / scale, gfd[c] / scale
        print(c, float(s[c]), gu, gfu)
        assert np.linalg.norm(gu - gfu) <= 2e-2 * np.linalg.norm(gu) + 0.05, (c, gu, gfu)


def test_autograd_function_matches_raw_adjoint():
    from loki_amd import refine

    torch = _torch()
    n, nbins, width = 1 << 16, 32, 4
    T = n * TSAMP
    ts_e, ts_v, t_ref = _inject(n, 41, [(44.0, 0.0, 0.0)], [0.08])
    th = np.array([[44.0 + 0.1 / T, 0.2 / T**2, 0.0], [44.0 - 0.05 / T, 0.0, 1.0 / T**3],
                   [44.01, 0.0, 0.0]])
    de, dv, dth = _dev(ts_e, ts_v, th)
    theta = dth.clone().requires_grad_(True)
    s = refine.fourier_score(theta, de, dv, TSAMP, t_ref, nbins, width)
    w = torch.tensor([1.0, -2.0, 0.5], device="cuda:0", dtype=torch.float32)
    (s * w).sum().backward()
    s_raw, st = refine.score_raw(de, dv, dth, TSAMP, t_ref, nbins, width)
    g_raw = refine.grad_raw(de, dv, dth, st, TSAMP, t_ref, nbins)
    torch.cuda.synchronize()
    assert torch.equal(s.detach(), s_raw)
    want = g_raw * w.double()[:, None]
    assert torch.allclose(theta.grad, want, rtol=0, atol=0)
    # forward without requires_grad skips the adjoint state; no graph is built
    s2 = refine.fourier_score(dth, de, dv, TSAMP, t_ref, nbins, width)
    assert not s2.requires_grad and torch.equal(s2, s_raw)


def test_refinement_climbs_to_injected_ephemeris():
    """Candidates started up to 0.4 / nbins turns off the injected (f, fdot, fddot) are driven
    uphill to one common maximum: the S/N never decreases, every candidate ends at the same S/N
    (at least that of the true ephemeris) and at the same point, which noise displaces from the
    truth by a small fraction of a turn."""
    from loki_amd import refine

    torch = _torch()
    n, nbins, width = 1 << 20, 32, 3
    T = n * TSAMP
    theta_true = (97.3, 3.0 / T**2, -6.0 / T**3)
    ts_e, ts_v, t_ref = _inject(n, 51, [theta_true], [0.03], fwhm=0.08)
    scale = np.array(refine.phase_scale(n, TSAMP, t_ref))
    rng = np.random.default_rng(52)
    du = rng.uniform(-0.4, 0.4, (16, 3)) / nbins
    th0 = np.asarray(theta_true)[None, :] + du / scale[None, :]
    de, dv, dth0 = _dev(ts_e, ts_v, th0)
    res = refine.refine_candidates(de, dv, dth0, TSAMP, t_ref, nbins, width, iters=40)
    s_true, _ = refine.score_raw(de, dv, torch.from_numpy(np.array([theta_true])).cuda(), TSAMP,
                                 t_ref, nbins, width, want_state=False)
    torch.cuda.synchronize()
    s0, s1 = res.scores0.cpu().numpy(), res.scores.cpu().numpy()
    st = float(s_true.cpu()[0])
    print("S true", st, "start", s0.round(2), "end", s1.round(2))
    assert np.all(s1 >= s0)
    assert np.all(s1 >= st - 1e-3)
    assert s1.max() - s1.min() < 0.02
    u1 = (res.theta.cpu().numpy() - np.asarray(theta_true)[None, :]) * scale[None, :]
    spread0 = np.linalg.norm(du - du.mean(axis=0), axis=1).max()
    spread1 = np.linalg.norm(u1 - u1.mean(axis=0), axis=1).max()
    print("spread (turns) start", spread0, "end", spread1, "offset", u1.mean(axis=0))
    # the maximum is a shallow ridge (f and fddot are nearly degenerate over a centred span):
    # candidates gather along it but need not meet in one point
    assert spread1 < 0.5 * spread0
    assert np.linalg.norm(u1.mean(axis=0)) < 0.1


def test_refine_errors():
    from loki_amd import refine

    torch = _torch()
    n = 4096
    de, dv = _dev(np.zeros(n, np.float32), np.ones(n, np.float32))
    th = torch.zeros((2, 3), dtype=torch.float64, device="cuda:0")
    th[:, 0] = 10.0
    with pytest.raises(ValueError, match="nbins"):
        refine.score_raw(de, dv, th, TSAMP, 0.0, 48, 4)
    with pytest.raises(ValueError, match="width"):
        refine.score_raw(de, dv, th, TSAMP, 0.0, 32, 32)
    with pytest.raises(ValueError, match="candidates"):
        refine.score_raw(de, dv, th[:0], TSAMP, 0.0, 32, 4)
    with pytest.raises(ValueError, match="float64"):
        refine.score_raw(de, dv, th.float(), TSAMP, 0.0, 32, 4)
    with pytest.raises(ValueError, match="shape"):
        refine.score_raw(de, dv, th[:, :2], TSAMP, 0.0, 32, 4)
    with pytest.raises(ValueError, match="tsamp"):
        refine.score_raw(de, dv, th, -1.0, 0.0, 32, 4)
    s, st = refine.score_raw(de, dv, th, TSAMP, 0.0, 32, 4)
    with pytest.raises(ValueError, match="state"):
        refine.grad_raw(de, dv, th, st[:1], TSAMP, 0.0, 32)
    ws = torch.empty(8, dtype=torch.uint8, device="cuda:0")
    from loki_amd.refine import _LIB

    rc = _LIB.loki_refine_score_device(de.data_ptr(), dv.data_ptr(), n, TSAMP, 0.0,
                                       th.data_ptr(), 2, 32, 4, s.data_ptr(), None,
                                       ws.data_ptr(), ws.numel(), None)
    assert rc == 1 and b"workspace" in _LIB.loki_last_error()
