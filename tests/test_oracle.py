"""Pinning the CPU oracle (CPU only).

The reference holds no golden vectors for fold / FFA / score (its tests compare implementations
at decimal=3), so the oracle is pinned by:
  * the current-version plan known answers logged in notebooks/loki_ep_{accel,jerk,circular}.ipynb;
  * the known answers of tests/cpp/utils_t.cpp (diff_max, circular_prefix_sum) and of
    generate_box_width_trials (score.cpp:390-405, worked by hand);
  * closed forms and brute force: exact-rational (fractions) resolution of the shift tables,
    histogram brute fold, textbook FP64 boxcar matched filter, numpy's inverse real FFT,
    direct DFT, integer-shift Fourier/time merge identity;
  * end-to-end physics: an injected pulsar is recovered at its true grid cell.
A plausible mistake (a flipped sign of delta_t, a swapped tail/head, an off-by-one bin) fails one
of these.
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle.plan import (KCVAL, Plan, SearchConfig, generate_box_width_trials, phase_idx,
                         phase_idx_uint)


# ---------------------------------------------------------------------------
# Plan known answers (notebooks, current grid rule)
# ---------------------------------------------------------------------------
EP_KNOWN = [
    # notebooks/loki_ep_accel.ipynb:208-209
    (dict(nsamps=2**25, tsamp=64e-6, nbins=64, eta=1.0,
          param_limits=[(-2500.0, 2500.0), (138.74643398, 151.34310645)],
          bseg_brute=2**25 // 16384, bseg_ffa=2**25 // 128),
     [16384, 1, 53, 2, 33], [128, 3, 6763, 2, 33], 2.55),
    # notebooks/loki_ep_jerk.ipynb:208-209
    (dict(nsamps=2**25, tsamp=64e-6, nbins=64, eta=1.0,
          param_limits=[(-2.5, 2.5), (-2184.35456, 3184.35456), (141.063088, 144.950108)],
          bseg_brute=2**25 // 8192, bseg_ffa=2**25 // 128),
     [8192, 1, 1, 33, 2, 33], [128, 1, 3, 2087, 2, 33], 0.79),
    # notebooks/loki_ep_circular.ipynb:246-247
    (dict(nsamps=2**24, tsamp=64e-6, nbins=64, eta=1.0,
          param_limits=[(-2.15432852e-03, 2.15432852e-03), (-3.68156042e-01, 3.68156042e-01),
                        (-6.29146716e+01, 6.29146716e+01), (-1.07515712e+04, 1.07515712e+04),
                        (1.41626654e+02, 1.44379474e+02)],
          bseg_brute=2**24 // 8192, bseg_ffa=2**24 // 128),
     [8192, 1, 1, 1, 1, 12, 2, 33], [128, 1, 1, 1, 3, 739, 2, 33], 0.28),
]


@pytest.mark.parametrize(("kw", "first", "last", "gb"), EP_KNOWN)
def test_plan_known_answers_from_notebooks(kw, first, last, gb):
    cfg = SearchConfig(use_fourier=True, **kw)
    plan = Plan(cfg, fourier=True)
    assert plan.fold_shapes[0] == first
    assert plan.fold_shapes[-1] == last
    assert round(plan.buffer_memory_usage, 2) == gb


def test_stale_notebooks_follow_floor_rule_not_ceil():
    """notebooks/loki_ffa_freq.ipynb:56-57 logged [16384, 86, 2, 64] -> [1, 1409630, 2, 64]
    under an older floor() grid rule; the current ceil() rule gives one more cell per axis."""
    cfg = SearchConfig(2**24, 64e-6, 64, 1.0, [(1 / 0.0075, 1 / 0.0065)],
                       bseg_brute=2**24 // 16384, use_fourier=False)
    plan = Plan(cfg)
    assert plan.fold_shapes[0] == [16384, 87, 2, 64]
    assert plan.fold_shapes[-1] == [1, 1409631, 2, 64]


def test_default_bseg_brute_and_levels():
    """configs.cpp:295-300: nsamps / 2^(floor(log2(tobs * f_min)) - init_levels)."""
    c1 = SearchConfig(2**20, 64e-6, 64, 1.0, [(0.5, 2.0)], use_fourier=False)
    assert c1.bseg_brute == 2**20 // 2**(5 - 1) and c1.niters_ffa == 4
    c2 = SearchConfig(2**23, 64e-6, 32, 1.0, [(-50, 50), (100, 400)], use_fourier=False)
    assert c2.bseg_brute == 2**23 // 2**(15 - 2) and c2.niters_ffa == 13


def test_config_validation_errors():
    with pytest.raises(ValueError):
        SearchConfig(1000, 64e-6, 64, 1.0, [(1.0, 2.0)], bseg_brute=8)
    with pytest.raises(ValueError):
        SearchConfig(2**16, 64e-6, 64, 1.0, [(1.0, 2.0)], bseg_brute=3)
    with pytest.raises(ValueError):
        SearchConfig(2**16, 64e-6, 64, 1.0, [(2.0, 1.0)], bseg_brute=1024)
    with pytest.raises(ValueError):
        SearchConfig(2**16, 64e-6, 64, 1.0, [(1.0, 2.0)], bseg_brute=2**16)


# ---------------------------------------------------------------------------
# Known answers of the reference's C++ unit tests and width trials
# ---------------------------------------------------------------------------
def test_diff_max_known_answers(oracle):
    """tests/cpp/utils_t.cpp:29-50."""
    assert oracle.diff_max([1, 2, 3, 4, 5], [0.5, 1.5, 2.5, 3.5, 4.5]) == 0.5
    assert oracle.diff_max([1, 4, 3, 2, 10], [0.5] * 5) == 9.5
    assert oracle.diff_max([1, 4, 3, 2, 10], [20] * 5) == -10.0


def test_circular_prefix_sum_known_answers(oracle):
    """tests/cpp/utils_t.cpp:52-79."""
    x = np.array([1, 2, 3, 4, 5], np.float32)
    np.testing.assert_array_equal(oracle.circular_prefix_sum(x, 5), [1, 3, 6, 10, 15])
    np.testing.assert_array_equal(oracle.circular_prefix_sum(x, 10),
                                  [1, 3, 6, 10, 15, 16, 18, 21, 25, 30])
    np.testing.assert_array_equal(oracle.circular_prefix_sum(x, 3), [1, 3, 6])


def test_box_width_trials_known_answers():
    """score.cpp:390-405 worked by hand: wmax = int(max(1, ducy*nbins)), next =
    max(w+1, int(wtsp*w))."""
    assert list(generate_box_width_trials(64, 0.2, 1.5)) == [1, 2, 3, 4, 6, 9]
    assert list(generate_box_width_trials(64, 0.3, 1.5)) == [1, 2, 3, 4, 6, 9, 13, 19]
    assert list(generate_box_width_trials(32, 0.5, 2.0)) == [1, 2, 4, 8, 16]
    assert list(generate_box_width_trials(8, 0.01, 1.5)) == [1]


# ---------------------------------------------------------------------------
# Phase / shift tables: closed forms and exact rationals
# ---------------------------------------------------------------------------
def test_phase_bin_closed_form():
    """tsamp = 1/64 s, f = 1 Hz, nbins = 64: sample k sits exactly in phase bin k mod 64."""
    k = np.arange(1000)
    np.testing.assert_array_equal(phase_idx_uint(k / 64.0, 1.0, 64), k % 64)
    # half-way phases round half up (psr_utils.hpp:51-63)
    assert phase_idx_uint(0.5 / 64.0, 1.0, 64) == 1
    # just below nbins wraps to bin 0
    assert phase_idx_uint(63.6 / 64.0, 1.0, 64) == 0


def _frac_phase(t, f, delay, nbins):
    x = (Fraction(t) - Fraction(delay)) * Fraction(f)
    return float((x - math.floor(x)) * nbins)


@pytest.mark.parametrize("lvl", [1, 2, 3])
def test_freq_resolver_against_exact_rationals(lvl):
    """taylor_ffa.cpp:249-271: idx = parent cell containing f, shift = frac(tseg_prev*f)*nbins."""
    cfg = SearchConfig(2**16, 64e-6, 64, 1.0, [(2.0, 8.0)], use_fourier=False)
    plan = Plan(cfg)
    if lvl >= plan.n_levels:
        pytest.skip("level beyond plan")
    got = plan.resolve_level_freq(lvl)
    lo, hi = map(Fraction, cfg.param_limits[0])
    n_cur, n_prev = plan.param_counts[lvl][0], plan.param_counts[lvl - 1][0]
    tseg = Fraction(cfg.tseg_brute) * 2 ** (lvl - 1)
    for i in range(0, n_cur, max(1, n_cur // 97)):
        f = lo + (hi - lo) / n_cur * (Fraction(i) + Fraction(1, 2))
        cell = math.floor(n_prev * (f - lo) / (hi - lo))
        assert got["idx"][i] == min(max(cell, 0), n_prev - 1)
        want = np.float32(_frac_phase(tseg, f, 0, 64))
        assert abs(float(got["shift"][i]) - float(want)) <= 2e-6 * 64


def _taylor_shift(j, a, dt):
    """transforms.cpp:101-120 (shift_taylor_params_d_f) in exact rationals: the kinematic
    Taylor vector (j, a, v=0, d=0) shifted by dt."""
    a_new = a + j * dt
    v_new = a * dt + j * dt * dt / 2
    d_new = a * dt * dt / 2 + j * dt * dt * dt / 6
    return a_new, v_new, d_new


@pytest.mark.parametrize("nparams", [2, 3])
def test_poly_resolver_against_exact_taylor_shift(nparams):
    """taylor_ffa.cpp:19-188 vs an independent exact-rational derivation (the generic path
    ffa_taylor_resolve_generic + shift_taylor_params_d_f, taylor_ffa.cpp:215-247): tail at
    delta_t = -tseg_prev/2, head at +tseg_prev/2; Doppler f' = f (1 - v/c); delay d/c."""
    if nparams == 2:
        lims = [(-60000.0, 60000.0), (100.0, 110.0)]
    else:
        lims = [(-2e5, 2e5), (-6e4, 6e4), (50.0, 55.0)]
    cfg = SearchConfig(2**17, 64e-6, 32, 1.0, lims, use_fourier=False)
    plan = Plan(cfg)
    c = Fraction(KCVAL)
    L = [(Fraction(a), Fraction(b)) for a, b in cfg.param_limits]
    checked = 0
    for lvl in range(1, plan.n_levels):
        tab = plan.resolve_level_poly(lvl)
        cur, prev = plan.param_counts[lvl], plan.param_counts[lvl - 1]
        tseg = Fraction(cfg.tseg_brute) * 2 ** (lvl - 1)
        strides = plan.param_cart_strides(lvl)
        for i in range(0, tab.size, max(1, tab.size // 23)):
            rem, idx = i, []
            for s in strides:
                idx.append(rem // s)
                rem %= s
            vals = [lo + (hi - lo) / n * (Fraction(k) + Fraction(1, 2)) if n > 1 else (lo + hi) / 2
                    for (lo, hi), n, k in zip(L, cur, idx)]
            j = vals[0] if nparams == 3 else Fraction(0)
            a, f = vals[-2], vals[-1]
            for side, (ifield, sfield) in ((-1, ("i_tail", "shift_tail")),
                                           (1, ("i_head", "shift_head"))):
                dt = side * tseg / 2
                a_new, v_new, d_new = _taylor_shift(j, a, dt)
                f_new = f * (1 - v_new / c)
                pnew = ([j] if nparams == 3 else []) + [a_new, f_new]
                flat = 0
                for (lo, hi), n, pv, s in zip(L, prev, pnew, Plan(cfg).param_cart_strides(lvl - 1)):
                    k = 0 if n <= 1 else min(max(math.floor(n * (pv - lo) / (hi - lo)
                                                            + Fraction(1, 10**6)), 0), n - 1)
                    flat += k * s
                assert int(tab[ifield][i]) == flat, (lvl, i, side)
                want = _frac_phase(dt, f, d_new / c, 32)
                got = float(tab[sfield][i])
                # circular distance in bins (a phase near 32 wraps to 0)
                dist = min(abs(got - want), 32 - abs(got - want))
                assert dist <= 1e-4, (lvl, i, side, got, want)
                checked += 1
    assert checked > 50


def test_resolver_invariants():
    cfg = SearchConfig(2**20, 64e-6, 64, 1.0, [(-1000, 1000), (100, 102)], use_fourier=False)
    plan = Plan(cfg)
    for lvl in range(1, plan.n_levels):
        tab = plan.resolve_level_poly(lvl)
        assert (tab["i_tail"] < plan.ncoords[lvl - 1]).all()
        assert (tab["i_head"] < plan.ncoords[lvl - 1]).all()
        for f in ("shift_tail", "shift_head"):
            assert (tab[f] >= 0).all() and (tab[f] < 64).all()


# ---------------------------------------------------------------------------
# Brute fold
# ---------------------------------------------------------------------------
def test_brute_fold_matches_histogram(oracle):
    rng = np.random.default_rng(0)
    ts_e = rng.standard_normal(2**14).astype(np.float32)
    ts_v = np.ones(2**14, np.float32)
    freqs = np.array([3.0, 7.7, 11.25])
    seglen, nbins, tsamp, t_ref = 4096, 16, 1e-3, 0.37
    fold = oracle.brute_fold_time(ts_e, ts_v, freqs, seglen, nbins, tsamp, t_ref)
    fold = fold.reshape(4, 3, 2, nbins)
    for s in range(4):
        for i, f in enumerate(freqs):
            k = np.arange(seglen)
            b = phase_idx_uint(k * tsamp - t_ref, f, nbins)
            e = np.zeros(nbins)
            np.add.at(e, b, ts_e[s * seglen:(s + 1) * seglen].astype(np.float64))
            cnt = np.bincount(b, minlength=nbins)
            np.testing.assert_allclose(fold[s, i, 0], e, rtol=1e-5, atol=1e-4)
            np.testing.assert_array_equal(fold[s, i, 1], cnt)  # ts_v = 1: exact counts


def test_brute_fold_sequential_order(oracle):
    """Each bucket is a sequential float32 sum in ascending sample order (kernels.cpp:569-574)."""
    rng = np.random.default_rng(1)
    ts_e = (rng.standard_normal(512) * 1e3).astype(np.float32)
    ts_v = rng.random(512).astype(np.float32)
    fold = oracle.brute_fold_time(ts_e, ts_v, np.array([13.0]), 512, 8, 1e-3, 0.0)
    b = phase_idx_uint(np.arange(512) * 1e-3, 13.0, 8)
    for q, x in ((0, ts_e), (1, ts_v)):
        for j in range(8):
            acc = np.float32(0)
            for k in np.nonzero(b == j)[0]:
                acc = np.float32(acc + x[k])
            assert fold[q * 8 + j] == acc


# ---------------------------------------------------------------------------
# Merges: physics and the Fourier/time identity
# ---------------------------------------------------------------------------
def test_ffa_recovers_injected_pulsar(oracle):
    """A noise-free pulse at an exact final-grid frequency concentrates in that cell: the FFA
    fold there has (nearly) the S/N of the direct full-length fold, and it is the maximum."""
    from loki_amd.synth import gaussian_profile, pulse_phase

    cfg = SearchConfig(2**16, 64e-6, 32, 1.0, [(2.0, 8.0)], use_fourier=False)
    plan = Plan(cfg)
    grid = plan.params_dict["freq"]
    f0 = grid[len(grid) // 3]
    ph = pulse_phase(cfg.nsamps, cfg.tsamp, f0)
    ts_e = (10.0 * gaussian_profile(ph, 0.05)).astype(np.float32)
    ts_v = np.ones(cfg.nsamps, np.float32)
    scores, _ = oracle.compute_ffa_scores(ts_e, ts_v, cfg)
    snr = scores.reshape(len(grid), -1).max(axis=1)
    assert abs(int(np.argmax(snr)) - len(grid) // 3) <= 1
    direct = oracle.brute_fold_time(ts_e, ts_v, np.array([f0]), cfg.nsamps, 32, cfg.tsamp, 0.0)
    d_snr = oracle.snr_boxcar_3d(direct, cfg.score_widths, nbins=32).max()
    assert snr.max() >= 0.9 * d_snr


def test_ffa_accel_recovers_injected_binary(oracle):
    """Accelerated pulse (reference time = mid-observation): best cell near the true (a, f)."""
    from loki_amd.synth import gaussian_profile, pulse_phase

    cfg = SearchConfig(2**18, 64e-6, 32, 1.0, [(-6000.0, 6000.0), (30.0, 34.0)],
                       use_fourier=False)
    plan = Plan(cfg)
    pd = plan.params_dict
    a0, f0 = 2500.0, 32.1
    ph = pulse_phase(cfg.nsamps, cfg.tsamp, f0, accel=a0)
    ts_e = (5.0 * gaussian_profile(ph, 0.05)).astype(np.float32)
    ts_v = np.ones(cfg.nsamps, np.float32)
    scores, _ = oracle.compute_ffa_scores(ts_e, ts_v, cfg)
    best = np.unravel_index(np.argmax(scores.reshape(len(pd["accel"]), len(pd["freq"]), -1)),
                            (len(pd["accel"]), len(pd["freq"]), cfg.score_widths.size))
    da = pd["accel"][1] - pd["accel"][0] if len(pd["accel"]) > 1 else 1e9
    df = pd["freq"][1] - pd["freq"][0]
    assert abs(pd["accel"][best[0]] - a0) <= 1.5 * da
    assert abs(pd["freq"][best[1]] - f0) <= 2 * df


def test_fourier_merge_equals_time_merge_for_integer_shifts(oracle):
    """With integer shifts, irfft(Fourier merge(rfft(tail), rfft(head))) equals the time-domain
    merge: pins the rotation direction and tail/head roles of kernels.cpp:71-146 / 241-453."""
    from oracle.plan import FFA_COORD_DTYPE

    rng = np.random.default_rng(3)
    nbins, nprev, ncur, nseg = 16, 5, 7, 2
    fold_in = rng.standard_normal((2 * nseg, nprev, 2, nbins)).astype(np.float32)
    coords = np.zeros(ncur, FFA_COORD_DTYPE)
    coords["i_tail"] = rng.integers(0, nprev, ncur)
    coords["i_head"] = rng.integers(0, nprev, ncur)
    coords["shift_tail"] = rng.integers(0, nbins, ncur).astype(np.float32)
    coords["shift_head"] = rng.integers(0, nbins, ncur).astype(np.float32)
    t_out = oracle.ffa_iter(fold_in.reshape(-1), coords, ncur, nprev, nseg, nbins)
    f_in = np.fft.rfft(fold_in.astype(np.float64), axis=-1).astype(np.complex64)
    f_out = oracle.ffa_iter(f_in.reshape(-1), coords, ncur, nprev, nseg, nbins, fourier=True)
    back = oracle.irfft(f_out, nbins)
    np.testing.assert_allclose(back, t_out, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("freq_only", [False, True])
def test_fourier_merge_fractional_shift_theorem(oracle, freq_only):
    """Fractional shifts: for band-limited rows x_j = sum_m a_m cos(2 pi m j / n + phi_m)
    (m < n/2), the merge equals rfft of x_tail(j - s_t) + x_head(j - s_h) evaluated
    analytically (the shift theorem), so a rounded shift, a sign flip or n vs n/2+1 in the
    twiddle fails here even where the integer-shift identity holds."""
    from oracle.plan import FFA_COORD_DTYPE, FFA_COORD_FREQ_DTYPE

    rng = np.random.default_rng(11)
    nbins, nprev, ncur, nseg = 32, 3, 6, 2
    j = np.arange(nbins)
    amp = rng.standard_normal((2 * nseg, nprev, 2, 15))
    phi = rng.uniform(0, 2 * np.pi, (2 * nseg, nprev, 2, 15))
    m = np.arange(1, 16)

    def rows(shift):
        arg = 2 * np.pi * m[:, None] * (j[None, :] - shift) / nbins
        return np.einsum("...m,...mj->...j", amp, np.cos(arg[None, None, None] + phi[..., None]))

    x = rows(0.0)  # [2*nseg, nprev, 2, nbins]
    f_in = np.fft.rfft(x, axis=-1).astype(np.complex64)
    if freq_only:
        coords = np.zeros(ncur, FFA_COORD_FREQ_DTYPE)
        coords["idx"] = rng.integers(0, nprev, ncur)
        coords["shift"] = rng.uniform(0, nbins, ncur).astype(np.float32)
        it, ih = coords["idx"], coords["idx"]
        st, sh = np.zeros(ncur), coords["shift"].astype(np.float64)
    else:
        coords = np.zeros(ncur, FFA_COORD_DTYPE)
        coords["i_tail"] = rng.integers(0, nprev, ncur)
        coords["i_head"] = rng.integers(0, nprev, ncur)
        coords["shift_tail"] = rng.uniform(0, nbins, ncur).astype(np.float32)
        coords["shift_head"] = rng.uniform(0, nbins, ncur).astype(np.float32)
        it, ih = coords["i_tail"], coords["i_head"]
        st = coords["shift_tail"].astype(np.float64)
        sh = coords["shift_head"].astype(np.float64)
    got = oracle.ffa_iter(f_in.reshape(-1), coords, ncur, nprev, nseg, nbins, fourier=True)
    got = got.reshape(nseg, ncur, 2, nbins // 2 + 1)
    for s in range(nseg):
        for c in range(ncur):
            arg_t = 2 * np.pi * m[:, None] * (j[None, :] - st[c]) / nbins
            arg_h = 2 * np.pi * m[:, None] * (j[None, :] - sh[c]) / nbins
            xt = np.einsum("qm,qmj->qj", amp[2 * s, it[c]],
                           np.cos(arg_t[None] + phi[2 * s, it[c]][..., None]))
            xh = np.einsum("qm,qmj->qj", amp[2 * s + 1, ih[c]],
                           np.cos(arg_h[None] + phi[2 * s + 1, ih[c]][..., None]))
            want = np.fft.rfft(xt + xh, axis=-1)
            np.testing.assert_allclose(got[s, c], want, rtol=0, atol=2e-4)


# ---------------------------------------------------------------------------
# Fourier helpers against numpy / direct sums
# ---------------------------------------------------------------------------
def test_irfft_matches_numpy(oracle):
    rng = np.random.default_rng(4)
    for nbins in (8, 16, 64, 33):
        X = (rng.standard_normal((6, nbins // 2 + 1))
             + 1j * rng.standard_normal((6, nbins // 2 + 1))).astype(np.complex64)
        np.testing.assert_allclose(oracle.irfft(X, nbins).reshape(6, nbins),
                                   np.fft.irfft(X.astype(np.complex128), n=nbins, axis=-1),
                                   rtol=1e-5, atol=1e-5)


def test_fourier_brute_fold_matches_direct_dft(oracle):
    rng = np.random.default_rng(5)
    ts_e = rng.standard_normal(1024).astype(np.float32)
    ts_v = np.ones(1024, np.float32)
    freqs = np.array([5.5, 13.0])
    nbins, tsamp, t_ref = 16, 1e-3, 0.256
    got = oracle.brute_fold_fourier(ts_e, ts_v, freqs, 512, nbins, tsamp, t_ref)
    got = got.reshape(2, 2, 2, nbins // 2 + 1)
    t = np.arange(512) * tsamp - t_ref
    for s in range(2):
        for i, f in enumerate(freqs):
            m = np.arange(nbins // 2 + 1)[:, None]
            w = np.exp(-2j * np.pi * m * f * t[None, :])
            want = w @ ts_e[s * 512:(s + 1) * 512].astype(np.float64)
            np.testing.assert_allclose(got[s, i, 0], want, rtol=1e-5, atol=1e-3)


def test_fourier_fold_is_dft_of_exact_phase_fold(oracle):
    """The harmonic-m Fourier fold equals the DFT (e^{-2 pi i m phase}) of the samples at their
    exact phases: for samples at exact bin centres it equals rfft of the time fold."""
    nbins, tsamp = 16, 1.0 / 16.0   # f = 1 Hz: sample k has phase exactly k/16
    ts_e = np.random.default_rng(6).standard_normal(64).astype(np.float32)
    ts_v = np.ones(64, np.float32)
    tf = oracle.brute_fold_time(ts_e, ts_v, np.array([1.0]), 64, nbins, tsamp, 0.0)
    ff = oracle.brute_fold_fourier(ts_e, ts_v, np.array([1.0]), 64, nbins, tsamp, 0.0)
    np.testing.assert_allclose(ff.reshape(2, -1)[0], np.fft.rfft(tf.reshape(2, -1)[0]),
                               rtol=1e-5, atol=1e-4)


# ---------------------------------------------------------------------------
# Boxcar S/N against the textbook matched filter
# ---------------------------------------------------------------------------
def _textbook_snr(prof, w):
    """Zero-mean, unit-norm boxcar template dot product, maximised over circular shifts."""
    n = prof.size
    best = -np.inf
    h = math.sqrt((n - w) / (n * w))
    b = w * h / (n - w)
    for s in range(n):
        tmpl = np.full(n, -b)
        tmpl[(s + np.arange(w)) % n] = h
        assert abs(tmpl.sum()) < 1e-12 and abs((tmpl ** 2).sum() - 1) < 1e-12
        best = max(best, float(tmpl @ prof))
    return best


def test_snr_boxcar_matches_textbook_template(oracle):
    rng = np.random.default_rng(7)
    nbins = 32
    e = rng.standard_normal((20, 1, nbins)) * 3
    v = rng.random((20, 1, nbins)) * 4 + 0.5
    folds = np.concatenate([e, v], axis=1).astype(np.float32)
    widths = np.array([1, 2, 5, 13, 31], np.uint64)
    got = oracle.snr_boxcar_3d(folds, widths)
    for p in range(20):
        prof = folds[p, 0].astype(np.float64) / np.sqrt(folds[p, 1].astype(np.float64))
        for k, w in enumerate(widths):
            assert abs(got[p, k] - _textbook_snr(prof, int(w))) <= 1e-4 * max(1, abs(got[p, k]))


def test_snr_boxcar_impulse_and_flat(oracle):
    nbins = 16
    widths = np.array([1, 2, 4, 8], np.uint64)
    imp = np.zeros((1, 2, nbins), np.float32)
    imp[0, 0, 5] = 1.0
    imp[0, 1, :] = 1.0
    got = oracle.snr_boxcar_3d(imp, widths)[0]
    want = [math.sqrt((nbins - w) / (nbins * w)) for w in widths]  # S/N = h_w
    np.testing.assert_allclose(got, want, rtol=1e-6)
    flat = np.ones((1, 2, nbins), np.float32)
    np.testing.assert_allclose(oracle.snr_boxcar_3d(flat, widths)[0], 0.0, atol=1e-6)


def test_f32_and_f64_oracle_folds_agree(oracle):
    cfg = SearchConfig(2**16, 64e-6, 32, 1.0, [(2.0, 8.0)], use_fourier=False)
    rng = np.random.default_rng(8)
    ts_e = rng.standard_normal(cfg.nsamps).astype(np.float32)
    ts_v = np.ones(cfg.nsamps, np.float32)
    f32, _ = oracle.compute_ffa(ts_e, ts_v, cfg)
    f64, _ = oracle.compute_ffa(ts_e, ts_v, cfg, f64=True)
    np.testing.assert_allclose(f32, f64, rtol=1e-4, atol=1e-3)


# ---------------------------------------------------------------------------
# Reduced-precision storage emulation
# ---------------------------------------------------------------------------
def test_round_storage_known_values(oracle):
    """Round to nearest even at 11 (FP16) and 8 (BF16) significand bits."""
    x = np.array([1.0, 1.0 + 2.0**-11, 1.0 + 3 * 2.0**-11, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8,
                  65504.0, 65520.0, -2.5, 2.0**-24, 3.0e38], np.float32)
    f16 = oracle.round_storage(x, "f16")
    np.testing.assert_array_equal(
        f16[:8], np.array([1.0, 1.0, 1.0 + 2 * 2.0**-10, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8,
                           65504.0, np.inf, -2.5], np.float32))
    assert f16[8] == 2.0**-24          # smallest FP16 subnormal
    bf = oracle.round_storage(x, "bf16")
    np.testing.assert_array_equal(
        bf[:6], np.array([1.0, 1.0, 1.0, 1.0, 1.0 + 4 * 2.0**-8, 65536.0], np.float32))
    assert bf[7] == -2.5 and np.isfinite(bf[9])


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_scale_round_storage_c_equals_numpy(oracle, storage):
    """The OpenMP C rounding used on whole levels equals numpy's (float16 cast; RNE on the
    FP32 pattern) on 2^22 random bit patterns of every finite magnitude, the binary16
    overflow / subnormal / tie edges and power-of-two scales."""
    rng = np.random.default_rng(99)
    bits = rng.integers(0, 0x7F800000, size=2**22, dtype=np.uint32)
    bits |= rng.integers(0, 2, size=bits.size, dtype=np.uint32) << 31
    x = bits.view(np.float32)
    edges = np.array([0.0, -0.0, 65504.0, 65519.996, 65520.0, -65520.0, 2.0**-14, 2.0**-24,
                      2.0**-25, 1.5 * 2.0**-24, 2.5 * 2.0**-24, 1.0 + 2.0**-11,
                      1.0 + 3 * 2.0**-11, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, 3.0e38, -3.4e38,
                      np.inf, -np.inf], np.float32)
    x = np.concatenate([x, edges])
    finite_scaled = {}
    for scale in (1.0, 0.5, 2.0**-9, 2.0**12):
        with np.errstate(over="ignore"):
            want = oracle.round_storage(x * np.float32(scale), storage)
        got = oracle.scale_round_storage(x.copy(), scale, storage)
        np.testing.assert_array_equal(got, want)
        finite_scaled[scale] = int(np.isfinite(got).sum())
    assert min(finite_scaled.values()) > 0


def test_reduced_storage_merge_is_scaled_sum(oracle):
    """Level l of the emulated tree equals round((t + h) * 0.5) of the level below, and the
    returned last level equals the unrounded sum times 2^(k0 + L - 2)."""
    from oracle.plan import Plan, SearchConfig

    cfg = SearchConfig(2**14, 64e-6, 16, 1.0, [(8.0, 12.0)], use_fourier=False)
    rng = np.random.default_rng(5)
    ts_e = rng.standard_normal(2**14, dtype=np.float32)
    ts_v = np.ones(2**14, np.float32)
    plan = Plan(cfg)
    k0 = oracle.storage_k0(cfg, plan)
    assert k0 == int(np.log2(plan.segment_lens[0] // cfg.nbins))
    f16, _ = oracle.compute_ffa(ts_e, ts_v, cfg, storage="f16")
    f32, _ = oracle.compute_ffa(ts_e, ts_v, cfg)
    # FP16 storage perturbs the FP32 result by about 2^-12 relative per level
    assert np.max(np.abs(f16 - f32)) <= 2.0**-12 * plan.n_levels * np.max(np.abs(f32))
    assert not np.array_equal(f16, f32)


# ---------------------------------------------------------------------------
# Block floating-point storage (bfp16) emulation
# ---------------------------------------------------------------------------
def _bfp16_exact(row):
    """Independent restatement in exact rationals: E from math.frexp of the row maximum,
    m = round-half-even(x / 2^E) clamped to +-32767, value m * 2^E."""
    import math
    from fractions import Fraction

    mx = max(abs(float(v)) for v in row)
    if mx == 0.0 or mx < 2.0**-126:
        e = -126
    else:
        _, k = math.frexp(mx)           # mx = f 2^k, f in [0.5, 1): ex = k - 1
        e = max(k - 1 - 14, -126)
    out = []
    for v in row:
        m = round(Fraction(float(v)) / Fraction(2) ** e)   # ties to even
        m = max(-32767, min(32767, m))
        out.append(float(Fraction(m) * Fraction(2) ** e))
    return np.array(out, np.float32)


def test_bfp16_known_values(oracle):
    """Hand-worked rows: exponent from the row maximum, 15-bit mantissas, ties to even,
    clamping at +-32767, zero and subnormal rows."""
    r = np.array([3.0, 1e-5, -2.5, 0.0], np.float32)        # max 3: E = 1 - 14 = -13
    np.testing.assert_array_equal(oracle.round_bfp16(r, 4), [3.0, 0.0, -2.5, 0.0])
    assert int(oracle.bfp16_exponent(np.float32(3.0))) == -13
    u = 2.0**-14                                             # max 1: E = -14
    r = np.array([1.0, 0.5 * u, 1.5 * u, 2.5 * u], np.float32)
    np.testing.assert_array_equal(oracle.round_bfp16(r, 4), [1.0, 0.0, 2 * u, 2 * u])
    r = np.array([65535.5, -65535.5, 1.0, 3.0], np.float32)  # E = 1: 32767.75 -> 32767
    np.testing.assert_array_equal(oracle.round_bfp16(r, 4), [65534.0, -65534.0, 0.0, 4.0])
    np.testing.assert_array_equal(oracle.round_bfp16(np.zeros(4, np.float32), 4), np.zeros(4))
    sub = np.full(4, 1e-40, np.float32)                      # E = -126: flushes to zero
    np.testing.assert_array_equal(oracle.round_bfp16(sub, 4), np.zeros(4))
    # rows are independent: a large row does not change its neighbour's exponent
    two = np.array([1000.0, 1.0, 0.001, 0.5], np.float32)
    np.testing.assert_array_equal(oracle.round_bfp16(two, 2)[2:],
                                  oracle.round_bfp16(two[2:], 2))


def test_bfp16_equals_exact_rationals_and_c(oracle):
    """numpy and the OpenMP C rounding both equal the exact-rational restatement on random
    rows spanning 60 binades (including rows of mixed magnitudes and exact ties)."""
    rng = np.random.default_rng(17)
    nbins = 16
    rows = []
    for k in range(300):
        scale = 2.0 ** rng.integers(-30, 30)
        r = (rng.standard_normal(nbins) * scale).astype(np.float32)
        if k % 3 == 0:
            r[rng.integers(0, nbins)] *= np.float32(2.0**12)   # one dominant bin
        if k % 5 == 0:  # exact ties at the row's mantissa step
            e = int(oracle.bfp16_exponent(np.abs(r).max()))
            r[1] = np.float32((rng.integers(-1000, 1000) + 0.5) * 2.0**e)
        rows.append(r)
    x = np.concatenate(rows)
    want = np.concatenate([_bfp16_exact(r) for r in rows])
    np.testing.assert_array_equal(oracle.round_bfp16(x, nbins), want)
    got = oracle.round_bfp16_rows(x.copy(), nbins)
    np.testing.assert_array_equal(got, want)
    # 15 significant bits relative to the row maximum
    err = np.abs(want - x).reshape(-1, nbins).max(axis=1)
    assert np.all(err <= np.abs(x).reshape(-1, nbins).max(axis=1) * 2.0**-15)


def test_bfp16_tree_is_rounded_sums(oracle):
    """The bfp16 tree: level 0 unrounded FP32, each stored level the row-rounded FP32 sum of
    the level below, the last level the unrounded sum; much closer to FP32 than FP16 is."""
    from oracle.plan import Plan, SearchConfig

    cfg = SearchConfig(2**14, 64e-6, 16, 1.0, [(8.0, 12.0)], use_fourier=False)
    rng = np.random.default_rng(5)
    ts_e = rng.standard_normal(2**14, dtype=np.float32)
    ts_v = np.ones(2**14, np.float32)
    plan = Plan(cfg)
    bfp, _ = oracle.compute_ffa(ts_e, ts_v, cfg, storage="bfp16")
    # the same tree by hand from the oracle's pinned kernels
    coords = plan.resolve_coordinates()
    fold = oracle.brute_fold_time(ts_e, ts_v, plan.brute_fold_freqs(), plan.segment_lens[0],
                                  cfg.nbins, cfg.tsamp, plan.brute_fold_tref())
    for lvl in range(1, plan.n_levels):
        fold = oracle.ffa_iter(fold, coords[lvl], plan.ncoords[lvl], plan.ncoords[lvl - 1],
                               plan.nsegments[lvl], cfg.nbins)
        if lvl < plan.n_levels - 1:
            fold = oracle.round_bfp16(fold, cfg.nbins)
    np.testing.assert_array_equal(bfp, fold)
    f32, _ = oracle.compute_ffa(ts_e, ts_v, cfg)
    f16, _ = oracle.compute_ffa(ts_e, ts_v, cfg, storage="f16")
    d_bfp = np.max(np.abs(bfp - f32))
    assert 0 < d_bfp <= 2.0**-15 * plan.n_levels * np.max(np.abs(f32))
    assert d_bfp < np.max(np.abs(f16 - f32)) / 4
