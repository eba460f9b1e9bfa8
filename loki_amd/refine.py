"""Differentiable Fourier-domain candidate scoring and gradient-ascent refinement on gfx950.

The score of a candidate theta = (f [Hz], fdot [Hz/s], fddot [Hz/s^2]) is the fixed-width boxcar
S/N (score.cpp:95-172) of the whole series folded in the Fourier domain at the continuous
ephemeris phi(t) = f tau + fdot tau^2 / 2 + fddot tau^3 / 6, tau = t - t_ref (the Fourier brute
fold of kernels_cuda.cu:780-853 over one segment spanning the series, the Taylor phase model of
transforms.cpp:101-120) and returned to nbins phase bins with the 1/n inverse real DFT
(fft_cuda.cu:357-448).  ``FourierScore`` is a ``torch.autograd.Function`` whose backward is the
hand-written adjoint kernel: it recomputes the fold from the series instead of storing it and
holds the boxcar position at the forward's maximum.  The reference has no counterpart of this
module; grid candidates of its FFA (ffa.cpp:207-234) are where refinement starts.

Everything runs through the C-ABI of ``include/loki_amd.h`` (``loki_refine_*``) on torch's
current stream; there is no CPU path.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import LIBRARY_PATH

_VP = ctypes.c_void_p
_SZ = ctypes.c_size_t
_D = ctypes.c_double


def _load():
    lib = ctypes.CDLL(str(LIBRARY_PATH))
    lib.loki_refine_workspace_bytes.argtypes = [_SZ, _SZ, _SZ]
    lib.loki_refine_workspace_bytes.restype = _SZ
    lib.loki_refine_score_device.argtypes = [_VP, _VP, _SZ, _D, _D, _VP, _SZ, _SZ, _SZ, _VP, _VP,
                                             _VP, _SZ, _VP]
    lib.loki_refine_score_device.restype = ctypes.c_int
    lib.loki_refine_grad_device.argtypes = [_VP, _VP, _SZ, _D, _D, _VP, _SZ, _SZ, _VP, _VP, _VP,
                                            _SZ, _VP]
    lib.loki_refine_grad_device.restype = ctypes.c_int
    lib.loki_last_error.restype = ctypes.c_char_p
    return lib


_LIB = _load()


def _check(rc: int) -> None:
    if rc != 0:
        msg = _LIB.loki_last_error().decode()
        if rc == 1:
            raise ValueError(msg)
        raise RuntimeError(msg)


def _stream(dev: torch.device) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def _series(ts_e: torch.Tensor, ts_v: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    if not (ts_e.is_cuda and ts_v.is_cuda):
        raise ValueError("ts_e and ts_v must be device tensors")
    if ts_e.dtype != torch.float32 or ts_v.dtype != torch.float32:
        raise ValueError("ts_e and ts_v must be float32")
    if ts_e.dim() != 1 or ts_e.shape != ts_v.shape:
        raise ValueError("ts_e and ts_v must be 1-D of equal length")
    return ts_e.contiguous(), ts_v.contiguous()


def _theta(theta: torch.Tensor, dev: torch.device) -> torch.Tensor:
    if theta.dim() != 2 or theta.shape[1] != 3:
        raise ValueError("theta must have shape [ncand, 3] = (f, fdot, fddot)")
    if theta.dtype != torch.float64:
        raise ValueError("theta must be float64")
    if theta.device != dev:
        raise ValueError("theta must be on the series' device")
    return theta.detach().contiguous()


def workspace_bytes(nsamps: int, ncand: int, nbins: int) -> int:
    return int(_LIB.loki_refine_workspace_bytes(nsamps, ncand, nbins))


def score_raw(ts_e: torch.Tensor, ts_v: torch.Tensor, theta: torch.Tensor, tsamp: float,
              t_ref: float, nbins: int, width: int, want_state: bool = True,
              workspace: torch.Tensor | None = None):
    """(scores [ncand] float32, adjoint state [ncand, 2, nbins/2+1] complex64 or None)."""
    ts_e, ts_v = _series(ts_e, ts_v)
    dev = ts_e.device
    th = _theta(theta, dev)
    ncand = th.shape[0]
    nws = workspace_bytes(ts_e.numel(), ncand, nbins)
    if workspace is None or workspace.numel() < nws:
        workspace = torch.empty(max(nws, 1), dtype=torch.uint8, device=dev)
    scores = torch.empty(ncand, dtype=torch.float32, device=dev)
    state = (torch.empty((ncand, 2, nbins // 2 + 1), dtype=torch.complex64, device=dev)
             if want_state else None)
    _check(_LIB.loki_refine_score_device(
        ts_e.data_ptr(), ts_v.data_ptr(), ts_e.numel(), float(tsamp), float(t_ref),
        th.data_ptr(), ncand, nbins, width, scores.data_ptr(),
        state.data_ptr() if state is not None else None, workspace.data_ptr(), workspace.numel(),
        _stream(dev)))
    return scores, state


def grad_raw(ts_e: torch.Tensor, ts_v: torch.Tensor, theta: torch.Tensor, state: torch.Tensor,
             tsamp: float, t_ref: float, nbins: int,
             workspace: torch.Tensor | None = None) -> torch.Tensor:
    """dS/dtheta [ncand, 3] float64 from the state of score_raw at the same theta."""
    ts_e, ts_v = _series(ts_e, ts_v)
    dev = ts_e.device
    th = _theta(theta, dev)
    ncand = th.shape[0]
    if state is None or state.shape != (ncand, 2, nbins // 2 + 1) or \
            state.dtype != torch.complex64 or not state.is_contiguous():
        raise ValueError("state must be the contiguous [ncand, 2, nbins/2+1] complex64 output "
                         "of score_raw")
    nws = workspace_bytes(ts_e.numel(), ncand, nbins)
    if workspace is None or workspace.numel() < nws:
        workspace = torch.empty(max(nws, 1), dtype=torch.uint8, device=dev)
    grad = torch.empty((ncand, 3), dtype=torch.float64, device=dev)
    _check(_LIB.loki_refine_grad_device(
        ts_e.data_ptr(), ts_v.data_ptr(), ts_e.numel(), float(tsamp), float(t_ref),
        th.data_ptr(), ncand, nbins, state.data_ptr(), grad.data_ptr(), workspace.data_ptr(),
        workspace.numel(), _stream(dev)))
    return grad


class FourierScore(torch.autograd.Function):
    """S(theta) for every candidate row of theta; differentiable in theta only."""

    @staticmethod
    def forward(ctx, theta, ts_e, ts_v, tsamp, t_ref, nbins, width):
        scores, state = score_raw(ts_e, ts_v, theta, tsamp, t_ref, nbins, width,
                                  want_state=theta.requires_grad)
        ctx.save_for_backward(theta, ts_e, ts_v, state)
        ctx.params = (float(tsamp), float(t_ref), int(nbins))
        return scores

    @staticmethod
    def backward(ctx, grad_scores):
        theta, ts_e, ts_v, state = ctx.saved_tensors
        tsamp, t_ref, nbins = ctx.params
        if not ctx.needs_input_grad[0]:
            return (None,) * 7
        g = grad_raw(ts_e, ts_v, theta, state, tsamp, t_ref, nbins)
        return (grad_scores.to(torch.float64)[:, None] * g, None, None, None, None, None, None)


def fourier_score(theta: torch.Tensor, ts_e: torch.Tensor, ts_v: torch.Tensor, tsamp: float,
                  t_ref: float, nbins: int, width: int) -> torch.Tensor:
    """Differentiable scores [ncand] (float32) of candidates theta [ncand, 3] (float64)."""
    return FourierScore.apply(theta, ts_e, ts_v, tsamp, t_ref, nbins, width)


def phase_scale(nsamps: int, tsamp: float, t_ref: float) -> tuple[float, float, float]:
    """Turns of phase a unit change of (f, fdot, fddot) moves the furthest sample by."""
    tmax = max(abs(-t_ref), abs((nsamps - 1) * tsamp - t_ref))
    return tmax, tmax * tmax / 2.0, tmax ** 3 / 6.0


@dataclass
class RefineResult:
    theta: torch.Tensor      # [ncand, 3] float64, refined
    scores: torch.Tensor     # [ncand] float32 at theta
    scores0: torch.Tensor    # [ncand] float32 at the start
    iterations: int
    evaluations: int         # forward + adjoint passes over the whole candidate set


def refine_candidates(ts_e: torch.Tensor, ts_v: torch.Tensor, theta0: torch.Tensor,
                      tsamp: float, t_ref: float, nbins: int, width: int, iters: int = 16,
                      step_turns: float | None = None, min_step_turns: float = 1e-4,
                      ) -> RefineResult:
    """Gradient ascent of the S/N with a per-candidate trust region.

    Parameters move in turn units u_i = theta_i * phase_scale_i (one unit = one turn of phase
    at the series' far end).  Each iteration tries u + r g_u / |g_u| for the candidate's step r,
    keeps it when the S/N rose (then r grows by 1.5) and otherwise halves r; candidates whose
    step fell below ``min_step_turns`` stop moving.  The S/N never decreases.
    """
    ts_e, ts_v = _series(ts_e, ts_v)
    dev = ts_e.device
    th = _theta(theta0, dev).clone()
    ncand = th.shape[0]
    scale = torch.tensor(phase_scale(ts_e.numel(), tsamp, t_ref), dtype=torch.float64,
                         device=dev)
    r = torch.full((ncand,), 0.25 / nbins if step_turns is None else step_turns,
                   dtype=torch.float64, device=dev)
    ws = torch.empty(max(workspace_bytes(ts_e.numel(), ncand, nbins), 1), dtype=torch.uint8,
                     device=dev)
    s_cur, state = score_raw(ts_e, ts_v, th, tsamp, t_ref, nbins, width, True, ws)
    g = grad_raw(ts_e, ts_v, th, state, tsamp, t_ref, nbins, ws)
    s0 = s_cur.clone()
    evals = 2
    it = 0
    for it in range(1, iters + 1):
        gu = g / scale
        norm = torch.linalg.vector_norm(gu, dim=1).clamp_min(1e-300)
        active = (r >= min_step_turns) & (norm > 1e-300)
        if not bool(active.any()):
            it -= 1
            break
        trial = th + torch.where(active, r / norm, 0.0)[:, None] * gu / scale
        s_try, st_try = score_raw(ts_e, ts_v, trial, tsamp, t_ref, nbins, width, True, ws)
        g_try = grad_raw(ts_e, ts_v, trial, st_try, tsamp, t_ref, nbins, ws)
        evals += 2
        acc = active & (s_try > s_cur)
        th = torch.where(acc[:, None], trial, th)
        s_cur = torch.where(acc, s_try, s_cur)
        g = torch.where(acc[:, None], g_try, g)
        r = torch.where(acc, r * 1.5, torch.where(active, r * 0.5, r))
    return RefineResult(theta=th, scores=s_cur, scores0=s0, iterations=it, evaluations=evals)


def accel_to_fdot(freq: float, accel: float, jerk: float = 0.0) -> tuple[float, float]:
    """(fdot, fddot) of a (freq, accel [m/s^2], jerk [m/s^3]) grid point.

    The observed frequency is f (1 - v / c) (transforms.cpp:100-121), so fdot = -f a / c and
    fddot = -f j / c to first order in v / c.
    """
    c = 299792458.0
    return -freq * accel / c, -freq * jerk / c


__all__ = ["FourierScore", "RefineResult", "accel_to_fdot", "fourier_score", "grad_raw",
           "phase_scale", "refine_candidates", "score_raw", "workspace_bytes"]
