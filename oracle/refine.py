"""FP64 reference of the differentiable Fourier-domain candidate score (test infrastructure).

For a candidate theta = (f, fdot, fddot) the series is folded over its full length in the
Fourier domain at the continuous ephemeris

    phi(t) = f tau + fdot tau^2 / 2 + fddot tau^3 / 6,   tau = t - t_ref,  t = k * tsamp,

    E_m = sum_k ts_e[k] exp(-2 pi i m phi(t_k)),   V_m likewise,   m = 0 .. nbins/2

(the Fourier brute fold of kernels.cpp:582-720 / kernels_cuda.cu:780-853 over one segment of the
whole series, with the phase model of transforms.cpp:101-120), returned to phase bins with the
1/n inverse real DFT (fft_cuda.cu:357-448), normalised p_j = e_j / sqrt(v_j) and scored with the
fixed-width boxcar of score.cpp:95-172:

    S = (h + b) * max_j W_j - b * sum_j p_j,   W_j = sum_{i < w} p_{(j + i) mod n},
    h = sqrt((n - w) / (n w)),  b = w h / (n - w).

The gradient dS/dtheta is exact away from ties of the window maximum (the maximising window j*
is held fixed).  Everything is FP64 numpy; it is the reference the GPU forward and adjoint
kernels are checked against, together with central finite differences of score().
"""

from __future__ import annotations

import numpy as np


def phases(nsamps: int, tsamp: float, t_ref: float, theta) -> tuple[np.ndarray, np.ndarray]:
    """(tau, phi) in FP64 for one candidate."""
    f, fd, fdd = (float(x) for x in theta)
    tau = np.arange(nsamps, dtype=np.float64) * tsamp - t_ref
    phi = f * tau + fd * tau * tau / 2.0 + fdd * tau * tau * tau / 6.0
    return tau, phi


def fold(ts_e, ts_v, tsamp: float, t_ref: float, theta, nbins: int):
    """Fourier fold (E, V), complex128 [nbins/2 + 1] each."""
    nh = nbins // 2 + 1
    _, phi = phases(len(ts_e), tsamp, t_ref, theta)
    fr = phi - np.floor(phi)
    m = np.arange(nh, dtype=np.float64)
    w = np.exp(-2j * np.pi * np.outer(m, fr))          # [nh, nsamps]
    return w @ np.asarray(ts_e, np.float64), w @ np.asarray(ts_v, np.float64)


def _irfft_matrix(nbins: int) -> np.ndarray:
    """Real linear map from (Re X_m, Im X_m) to x_j of numpy.fft.irfft (n = nbins)."""
    nh = nbins // 2 + 1
    eye = np.eye(nh)
    cols_re = np.fft.irfft(eye, n=nbins, axis=1)         # [nh, nbins]: d x / d Re X_m
    cols_im = np.fft.irfft(1j * eye, n=nbins, axis=1)    # d x / d Im X_m
    return cols_re, cols_im


def score_terms(E, V, nbins: int, width: int):
    """(S, e, v, p, jstar) from the Fourier fold."""
    e = np.fft.irfft(E, n=nbins)
    v = np.fft.irfft(V, n=nbins)
    p = e / np.sqrt(v)
    n, w = nbins, int(width)
    h = np.sqrt((n - w) / (n * w))
    b = w * h / (n - w)
    ext = np.concatenate([p, p[: w - 1]])
    win = np.array([ext[j : j + w].sum() for j in range(n)])
    jstar = int(np.argmax(win))
    s = (h + b) * win[jstar] - b * p.sum()
    return s, e, v, p, jstar


def score(ts_e, ts_v, tsamp: float, t_ref: float, theta, nbins: int, width: int) -> float:
    E, V = fold(ts_e, ts_v, tsamp, t_ref, theta, nbins)
    return float(score_terms(E, V, nbins, width)[0])


def grad(ts_e, ts_v, tsamp: float, t_ref: float, theta, nbins: int, width: int):
    """(S, dS/dtheta) by the chain rule (the adjoint the GPU kernels implement)."""
    nh = nbins // 2 + 1
    E, V = fold(ts_e, ts_v, tsamp, t_ref, theta, nbins)
    s, e, v, p, jstar = score_terms(E, V, nbins, width)
    n, w = nbins, int(width)
    h = np.sqrt((n - w) / (n * w))
    b = w * h / (n - w)
    dp = np.full(n, -b)
    dp[(jstar + np.arange(w)) % n] += h + b
    de = dp / np.sqrt(v)
    dv = -dp * e / (2.0 * v * np.sqrt(v))
    cre, cim = _irfft_matrix(nbins)
    # dS/dRe X_m, dS/dIm X_m
    gE = cre @ de + 1j * (cim @ de)
    gV = cre @ dv + 1j * (cim @ dv)
    tau, phi = phases(len(ts_e), tsamp, t_ref, theta)
    fr = phi - np.floor(phi)
    m = np.arange(nh, dtype=np.float64)
    w_mk = np.exp(-2j * np.pi * np.outer(m, fr))         # [nh, nsamps]
    # dE_m/dphi_k = x_k (-2 pi i m) w_mk ; dS/dphi_k = sum_m Re(conj(g_m) dE_m/dphi_k)
    dphi = (np.real(np.conj(gE)[:, None] * (-2j * np.pi * m)[:, None] * w_mk) *
            np.asarray(ts_e, np.float64)[None, :]).sum(axis=0)
    dphi += (np.real(np.conj(gV)[:, None] * (-2j * np.pi * m)[:, None] * w_mk) *
             np.asarray(ts_v, np.float64)[None, :]).sum(axis=0)
    g = np.array([np.sum(dphi * tau), np.sum(dphi * tau * tau / 2.0),
                  np.sum(dphi * tau * tau * tau / 6.0)])
    return float(s), g


def grad_fd(ts_e, ts_v, tsamp: float, t_ref: float, theta, nbins: int, width: int,
            rel: float = 1e-6, scale=None):
    """Central finite differences of score() with step rel * scale_i per parameter."""
    theta = np.asarray(theta, np.float64)
    scale = np.abs(theta) if scale is None else np.asarray(scale, np.float64)
    g = np.zeros(3)
    for i in range(3):
        hstep = rel * scale[i]
        tp, tm = theta.copy(), theta.copy()
        tp[i] += hstep
        tm[i] -= hstep
        g[i] = (score(ts_e, ts_v, tsamp, t_ref, tp, nbins, width)
                - score(ts_e, ts_v, tsamp, t_ref, tm, nbins, width)) / (2.0 * hstep)
    return g
