"""Seeded synthetic time series for tests and benchmarks.

``ts_e`` is N(0, 1) float32 noise (numpy PCG64 from a fixed seed) plus an optional periodic
Gaussian pulse; ``ts_v`` is 1.  The pulse phase follows the Taylor model the FFA grid searches,
referenced to the middle of the observation (the reference's t_ref = tseg/2 for nparams > 1):

    phi(t) = f0 * (tau - (a tau^2 / 2 + j tau^3 / 6) / c),   tau = t - tobs / 2   (nparams > 1)
    phi(t) = f0 * t                                                              (freq only)

The amplitude is solved so that the noise-free fold at the true parameters, scored with the
boxcar S/N formula (score.cpp:95-172) at the best width, reaches the requested S/N.
"""

from __future__ import annotations

import numpy as np

C_LIGHT = 299792458.0


def pulse_phase(nsamps: int, tsamp: float, f0: float, accel: float = 0.0, jerk: float = 0.0,
                t_ref: float | None = None) -> np.ndarray:
    t = np.arange(nsamps, dtype=np.float64) * tsamp
    if t_ref is None:
        t_ref = 0.5 * nsamps * tsamp if (accel != 0.0 or jerk != 0.0) else 0.0
    tau = t - t_ref
    d = 0.5 * accel * tau**2 + jerk * tau**3 / 6.0
    return f0 * (tau - d / C_LIGHT)


def gaussian_profile(phase: np.ndarray, fwhm: float) -> np.ndarray:
    """Periodic Gaussian of unit height centred at phase 0.5."""
    sigma = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    x = (phase - np.floor(phase)) - 0.5
    return np.exp(-0.5 * (x / sigma) ** 2)


def boxcar_snr_noise_free(profile: np.ndarray, count: np.ndarray, widths) -> float:
    """Boxcar S/N of a noise-free fold e = profile, v = count (FP64)."""
    n = profile.size
    p = profile / np.sqrt(count)
    ps = np.concatenate([np.cumsum(p), np.cumsum(p) + p.sum()])
    best = -np.inf
    for w in widths:
        if w >= n:
            continue
        h = np.sqrt((n - w) / (n * w))
        b = w * h / (n - w)
        d = np.max(ps[w:w + n] - ps[:n])
        best = max(best, (h + b) * d - b * p.sum())
    return float(best)


def make_series(nsamps: int, tsamp: float, seed: int, f0: float | None = None,
                accel: float = 0.0, jerk: float = 0.0, fwhm: float = 0.05, snr: float = 0.0,
                nbins: int = 64, widths=(1, 2, 3, 4, 6, 8, 12, 16, 24, 32),
                t_ref: float | None = None):
    """Return (ts_e, ts_v, amplitude)."""
    rng = np.random.default_rng(seed)
    ts_e = rng.standard_normal(nsamps, dtype=np.float32)
    ts_v = np.ones(nsamps, dtype=np.float32)
    amp = 0.0
    if f0 is not None and snr > 0:
        ph = pulse_phase(nsamps, tsamp, f0, accel, jerk, t_ref)
        shape = gaussian_profile(ph, fwhm)
        bins = np.floor((ph - np.floor(ph)) * nbins).astype(np.int64) % nbins
        prof = np.bincount(bins, weights=shape, minlength=nbins)
        cnt = np.bincount(bins, minlength=nbins).astype(np.float64)
        unit = boxcar_snr_noise_free(prof, np.maximum(cnt, 1.0), widths)
        amp = snr / unit if unit > 0 else 0.0
        ts_e += (amp * shape).astype(np.float32)
    return ts_e, ts_v, amp


def make_injections(n: int, seed: int, f_range=(50.0, 500.0), fdot_max: float = 1e-8,
                    fddot_max: float = 1e-14, snr_range=(10.0, 25.0), fwhm: float = 0.05):
    """n signals with f, fdot, fddot and S/N drawn uniformly (FP64), sorted by f."""
    rng = np.random.default_rng(seed)
    sig = [dict(f0=float(rng.uniform(*f_range)), fdot=float(rng.uniform(-fdot_max, fdot_max)),
                fddot=float(rng.uniform(-fddot_max, fddot_max)),
                snr=float(rng.uniform(*snr_range)), fwhm=fwhm) for _ in range(n)]
    return sorted(sig, key=lambda d: d["f0"])


def make_injected_series(nsamps: int, tsamp: float, seed: int, signals, nbins: int = 64,
                         widths=(1, 2, 3, 4, 6, 8, 12, 16, 24, 32), device=None):
    """N(0,1) noise plus several pulses with the continuous ephemeris
    phi = f0 tau + fdot tau^2 / 2 + fddot tau^3 / 6, tau = t - tobs / 2; each amplitude is
    solved for that signal's S/N as in make_series.  Returns (ts_e, ts_v, amplitudes) as numpy
    arrays.  With a torch ``device`` the FP64 phases and pulse shapes are evaluated there (same
    arithmetic; the noise always comes from numpy's PCG64)."""
    rng = np.random.default_rng(seed)
    ts_e = rng.standard_normal(nsamps, dtype=np.float32)
    ts_v = np.ones(nsamps, dtype=np.float32)
    amps = []
    if device is None:
        tau = np.arange(nsamps, dtype=np.float64) * tsamp - 0.5 * nsamps * tsamp
        acc = np.zeros(nsamps, np.float64)
        for sg in signals:
            ph = sg["f0"] * tau + sg["fdot"] * tau * tau / 2.0 + sg["fddot"] * tau**3 / 6.0
            shape = gaussian_profile(ph, sg["fwhm"])
            bins = np.floor((ph - np.floor(ph)) * nbins).astype(np.int64) % nbins
            prof = np.bincount(bins, weights=shape, minlength=nbins)
            cnt = np.bincount(bins, minlength=nbins).astype(np.float64)
            unit = boxcar_snr_noise_free(prof, np.maximum(cnt, 1.0), widths)
            amp = sg["snr"] / unit if unit > 0 else 0.0
            acc += (amp * shape).astype(np.float32)
            amps.append(amp)
        return (ts_e + acc).astype(np.float32), ts_v, amps
    import torch

    tau = torch.arange(nsamps, dtype=torch.float64, device=device) * tsamp - 0.5 * nsamps * tsamp
    acc = torch.zeros(nsamps, dtype=torch.float64, device=device)
    for sg in signals:
        ph = sg["f0"] * tau + sg["fdot"] * tau * tau / 2.0 + sg["fddot"] * tau**3 / 6.0
        fr = ph - torch.floor(ph)
        sigma = sg["fwhm"] / (2.0 * np.sqrt(2.0 * np.log(2.0)))
        shape = torch.exp(-0.5 * ((fr - 0.5) / sigma) ** 2)
        bins = torch.floor(fr * nbins).to(torch.int64) % nbins
        prof = torch.bincount(bins, weights=shape, minlength=nbins).cpu().numpy()
        cnt = torch.bincount(bins, minlength=nbins).to(torch.float64).cpu().numpy()
        unit = boxcar_snr_noise_free(prof, np.maximum(cnt, 1.0), widths)
        amp = sg["snr"] / unit if unit > 0 else 0.0
        acc += (amp * shape).to(torch.float32).to(torch.float64)
        amps.append(amp)
    return (ts_e + acc.cpu().numpy()).astype(np.float32), ts_v, amps


def make_batch_specs(nseries: int, seed: int, log2_len=(18, 22), nbins_choices=(32, 64, 128),
                     tsamp: float = 64e-6, period_centre=(0.3, 3.0), min_turns: float = 4.0):
    """Specs of a batch of independent period searches of uneven length.

    Series i (numpy PCG64 seeded with seed + i): nsamps = 2^round(U(log2_len)); nbins drawn
    from nbins_choices; a period centre P drawn uniformly in period_centre (capped so that the
    observation spans at least ``min_turns`` turns of the longest period, which keeps at least
    one FFA merge level), searched over periods [P/2, 2P], i.e. frequencies [1/(2P), 2/P];
    every other series carries a pulse with period uniform in [P/2, 2P], FWHM 2-10 % of the
    period and S/N 8-30.  Returns a list of dicts (nsamps, nbins, f_lo, f_hi, seed, signal).
    """
    specs = []
    for i in range(nseries):
        rng = np.random.default_rng(seed + i)
        nsamps = 2 ** int(round(rng.uniform(*log2_len)))
        nbins = int(rng.choice(nbins_choices))
        tobs = nsamps * tsamp
        p_hi = min(period_centre[1], tobs / (2.0 * min_turns))
        p_lo = min(period_centre[0], 0.5 * p_hi)
        p_c = rng.uniform(p_lo, p_hi)
        sig = None
        if i % 2 == 0:
            period = rng.uniform(0.5 * p_c, 2.0 * p_c)
            sig = dict(f0=1.0 / period, fwhm=float(rng.uniform(0.02, 0.10)),
                       snr=float(rng.uniform(8.0, 30.0)))
        specs.append(dict(nsamps=nsamps, nbins=nbins, f_lo=1.0 / (2.0 * p_c), f_hi=2.0 / p_c,
                          seed=seed + i, signal=sig))
    return specs


def make_batch_series(spec: dict, tsamp: float = 64e-6, widths=(1, 2, 3, 4, 6, 8, 12, 16, 24, 32)):
    """(ts_e, ts_v) of one batch spec (see make_batch_specs)."""
    sig = spec["signal"]
    w = [x for x in widths if x < spec["nbins"]]
    if sig is None:
        e, v, _ = make_series(spec["nsamps"], tsamp, spec["seed"])
    else:
        e, v, _ = make_series(spec["nsamps"], tsamp, spec["seed"], f0=sig["f0"], fwhm=sig["fwhm"],
                              snr=sig["snr"], nbins=spec["nbins"], widths=w)
    return e, v
