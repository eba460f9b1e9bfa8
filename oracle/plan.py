"""Oracle restatement of the FFA config, plan geometry and shift tables.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

All floating-point expressions are evaluated in IEEE binary64 in the same
operation order as the reference source, with no fused multiply-adds, so the
coordinate tables produced here are the bit-exact definition the HIP resolver
kernels are tested against.

Citations are to pravirkr/loki v0.0.3.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# lib/loki/utils.hpp:11-14
KCVAL = 299792458.0
KINVCVAL = 3.3356409519815204e-09
KFLOATEPS = 1e-6
# include/loki/common/types.hpp:87-88
PARAM_NAMES = ("crackle", "snap", "jerk", "accel", "freq")

FFA_COORD_DTYPE = np.dtype(
    [("i_tail", "<u4"), ("shift_tail", "<f4"), ("i_head", "<u4"), ("shift_head", "<f4")]
)  # include/loki/common/coord.hpp:15-20
FFA_COORD_FREQ_DTYPE = np.dtype([("idx", "<u4"), ("shift", "<f4")])  # coord.hpp:22-25


def is_power_of_2(n: int) -> bool:
    return n > 0 and (n & (n - 1)) == 0


# ---------------------------------------------------------------------------
# Grid helpers (lib/psr_utils.cpp:14-45, lib/loki/psr_utils.hpp:32-63,235-320)
# ---------------------------------------------------------------------------
def poly_taylor_step_f(nparams: int, tobs: float, fold_bins: int, tol_bins: float,
                       t_ref: float = 0.0) -> list[float]:
    """psr_utils.cpp:14-29: dparam_k = dphi * 2^i * (i+1)! / dt^(i+1)."""
    dphi = tol_bins / float(fold_bins)
    dt = tobs - t_ref
    out = [0.0] * nparams
    for i in range(nparams):
        orth = math.pow(2.0, i)
        out[nparams - 1 - i] = dphi * orth * math.gamma(float(i + 1) + 1.0) / math.pow(dt, i + 1)
    return out


def poly_taylor_step_d_f(nparams: int, tobs: float, fold_bins: int, tol_bins: float,
                         f_max: float, t_ref: float = 0.0) -> list[float]:
    """psr_utils.cpp:31-45: non-freq steps converted to physical units * c / f_max."""
    df = poly_taylor_step_f(nparams, tobs, fold_bins, tol_bins, t_ref)
    out = [0.0] * nparams
    for i in range(nparams - 1):
        out[i] = df[i] * KCVAL / f_max
    out[nparams - 1] = df[nparams - 1]
    return out


def range_param_count(vmin: float, vmax: float, dv: float) -> int:
    """psr_utils.hpp:235-242 (ceil rule)."""
    if not vmax > vmin:
        raise ValueError("vmax must be greater than vmin")
    if not dv > 0:
        raise ValueError("dv must be positive")
    if dv >= (vmax - vmin):
        return 1
    return int(math.ceil((vmax - vmin) / dv))


def range_param(vmin: float, vmax: float, dv: float) -> np.ndarray:
    """psr_utils.hpp:254-270: cell centres start + dv_actual * i."""
    if not vmax > vmin:
        raise ValueError("vmax must be greater than vmin")
    if not dv > 0:
        raise ValueError("dv must be positive")
    if dv >= (vmax - vmin):
        return np.array([(vmax + vmin) / 2.0])
    npoints = int(math.ceil((vmax - vmin) / dv))
    dv_actual = (vmax - vmin) / float(npoints)
    start = vmin + (0.5 * dv_actual)
    return start + dv_actual * np.arange(npoints, dtype=np.float64)


def param_val_at_idx(vmin: float, vmax: float, count: int, i):
    """psr_utils.hpp:283-293: vmin + dv_actual * (i + 0.5) (note: differs in
    rounding from range_param's start + dv*i)."""
    i = np.asarray(i, dtype=np.float64)
    if count <= 1:
        return np.full(i.shape, (vmax + vmin) / 2.0)
    dv_actual = (vmax - vmin) / float(count)
    return vmin + (dv_actual * (i + 0.5))


def nearest_idx_analytical(val, vmin: float, vmax: float, count: int) -> np.ndarray:
    """psr_utils.hpp:304-320: int(count*(val-min)/(max-min) + 1e-6), clamped."""
    val = np.asarray(val, dtype=np.float64)
    if count <= 1:
        return np.zeros(val.shape, dtype=np.int64)
    raw = (float(count) * (val - vmin)) / (vmax - vmin)
    idx = np.trunc(raw + KFLOATEPS)
    idx = np.clip(idx, 0, count - 1)
    return idx.astype(np.int64)


def phase_idx(proper_time, freq, nbins: int, delay=0.0) -> np.ndarray:
    """psr_utils.hpp:32-49: frac((t - delay) * f) * nbins -> float32 (0 if == nbins)."""
    total = (np.asarray(proper_time, dtype=np.float64) - delay) * freq
    norm = total - np.floor(total)
    iphase = norm * float(nbins)
    iphase = np.where(iphase >= float(nbins), 0.0, iphase)
    return iphase.astype(np.float32)


def phase_idx_uint(proper_time, freq, nbins: int, delay=0.0) -> np.ndarray:
    """psr_utils.hpp:53-63: u32(f32 phase + 0.5f), nbins -> 0."""
    ph = phase_idx(proper_time, freq, nbins, delay)
    b = (ph + np.float32(0.5)).astype(np.float32).astype(np.uint32)
    return np.where(b == nbins, 0, b).astype(np.uint32)


def generate_box_width_trials(fold_bins: int, ducy_max: float = 0.2, wtsp: float = 1.5) -> np.ndarray:
    """score.cpp:390-405."""
    wmax = int(max(1.0, ducy_max * float(fold_bins)))
    widths = [1]
    while widths[-1] < wmax:
        nxt = max(widths[-1] + 1, int(wtsp * float(widths[-1])))
        if nxt > wmax:
            break
        widths.append(nxt)
    return np.array(widths, dtype=np.uint64)


# ---------------------------------------------------------------------------
# PulsarSearchConfig (configs.hpp:22-164, configs.cpp:18-303)
# ---------------------------------------------------------------------------
@dataclass
class SearchConfig:
    nsamps: int
    tsamp: float
    nbins: int
    eta: float
    param_limits: list
    ducy_max: float = 0.2
    wtsp: float = 1.5
    use_fourier: bool = True
    bseg_brute: int | None = None
    bseg_ffa: int | None = None
    nbins_max: int = 1024
    nbins_min_lossy_bf: int = 64
    # derived
    nparams: int = field(init=False)
    nbins_f: int = field(init=False)
    f_min: float = field(init=False)
    f_max: float = field(init=False)
    tseg_brute: float = field(init=False)
    tseg_ffa: float = field(init=False)
    niters_ffa: int = field(init=False)
    score_widths: np.ndarray = field(init=False)
    param_names: tuple = field(init=False)

    def __post_init__(self) -> None:
        lim = [(float(a), float(b)) for a, b in np.asarray(self.param_limits, dtype=np.float64).reshape(-1, 2)]
        if not lim:
            raise ValueError("coord_limits must be non-empty")
        self.param_limits = lim
        self.nbins_f = self.nbins // 2 + 1
        self.nparams = len(lim)
        self.param_names = PARAM_NAMES[len(PARAM_NAMES) - self.nparams:]
        self.f_min, self.f_max = lim[-1]
        if self.bseg_brute is None:
            init_levels = 1 if self.nparams == 1 else 2
            levels = int(math.log2(float(self.nsamps) * self.tsamp * self.f_min))
            self.bseg_brute = self.nsamps // (1 << (levels - init_levels))
        if self.bseg_ffa is None:
            self.bseg_ffa = self.nsamps
        self._validate()
        self.tseg_brute = float(self.bseg_brute) * self.tsamp
        self.tseg_ffa = float(self.bseg_ffa) * self.tsamp
        self.niters_ffa = int(math.log2(float(self.bseg_ffa) / float(self.bseg_brute)))
        self.score_widths = generate_box_width_trials(self.nbins, self.ducy_max, self.wtsp)

    def _validate(self) -> None:  # configs.cpp:262-293
        if not is_power_of_2(self.nsamps):
            raise ValueError("nsamps must be a power of 2")
        if not self.tsamp > 0:
            raise ValueError("tsamp must be positive")
        if not self.eta > 0:
            raise ValueError("eta must be positive")
        if not self.nbins_max >= self.nbins:
            raise ValueError("nbins_max must be >= nbins")
        if not is_power_of_2(self.bseg_brute):
            raise ValueError("bseg_brute must be a power of 2")
        if not is_power_of_2(self.bseg_ffa):
            raise ValueError("bseg_ffa must be a power of 2")
        if not self.bseg_brute < self.nsamps:
            raise ValueError("bseg_brute must be less than nsamps")
        if not self.bseg_ffa <= self.nsamps:
            raise ValueError("bseg_ffa must be <= nsamps")
        if not self.bseg_ffa >= self.bseg_brute:
            raise ValueError("bseg_ffa must be >= bseg_brute")
        for lo, hi in self.param_limits:
            if not hi >= lo:
                raise ValueError("param_limits must be increasing")

    @property
    def tobs(self) -> float:
        return float(self.nsamps) * self.tsamp

    def dparams(self, tseg_cur: float) -> list[float]:  # configs.cpp:158-162
        t_ref = 0.0 if self.nparams == 1 else tseg_cur / 2.0
        return poly_taylor_step_d_f(self.nparams, tseg_cur, self.nbins, self.eta, self.f_max, t_ref)

    def param_grid_count(self, tseg_cur: float) -> list[int]:  # configs.cpp:175-184
        dp = self.dparams(tseg_cur)
        return [range_param_count(lo, hi, d) for (lo, hi), d in zip(self.param_limits, dp)]

    def dparams_actual(self, tseg_cur: float) -> list[float]:  # configs.cpp:164-174
        cnt = self.param_grid_count(tseg_cur)
        return [(hi - lo) / float(c) for (lo, hi), c in zip(self.param_limits, cnt)]


# ---------------------------------------------------------------------------
# FFAPlan (plans.cpp:264-343, 445-565)
# ---------------------------------------------------------------------------
class Plan:
    def __init__(self, cfg: SearchConfig, fourier: bool = False) -> None:
        self.cfg = cfg
        self.fourier = fourier
        levels = cfg.niters_ffa + 1
        self.n_levels = levels
        self.n_params = cfg.nparams
        self.segment_lens, self.nsegments, self.tsegments = [], [], []
        self.param_counts, self.dparams, self.dparams_act, self.ncoords = [], [], [], []
        for lvl in range(levels):
            seglen = cfg.bseg_brute * (1 << lvl)
            tseg = float(seglen) * cfg.tsamp
            self.segment_lens.append(seglen)
            self.nsegments.append(cfg.nsamps // seglen)
            self.tsegments.append(tseg)
            self.dparams.append(cfg.dparams(tseg))
            self.dparams_act.append(cfg.dparams_actual(tseg))
            cnt = cfg.param_grid_count(tseg)
            self.param_counts.append(cnt)
            self.ncoords.append(int(np.prod(cnt, dtype=object)))
        self.ncoords_offsets = [0]
        for n in self.ncoords:
            self.ncoords_offsets.append(self.ncoords_offsets[-1] + n)
        for i in range(self.n_params - 1):  # plans.cpp:322-333
            if self.param_counts[0][i] != 1:
                raise ValueError("Only one parameter for higher order derivatives is "
                                 "supported for the initialization for the first level")
        W = cfg.nbins_f if fourier else cfg.nbins
        self.fold_shapes = [[self.nsegments[l], *self.param_counts[l], 2, W] for l in range(levels)]
        self.fold_shapes_time = [[self.nsegments[l], *self.param_counts[l], 2, cfg.nbins]
                                 for l in range(levels)]

    @staticmethod
    def _prod(shape) -> int:
        p = 1
        for s in shape:
            p *= int(s)
        return p

    @property
    def brute_fold_size(self) -> int:
        return self._prod(self.fold_shapes[0])

    @property
    def fold_size(self) -> int:
        return self._prod(self.fold_shapes[-1])

    @property
    def fold_size_time(self) -> int:
        return self._prod(self.fold_shapes_time[-1])

    @property
    def buffer_size(self) -> int:
        return max(self._prod(s) for s in self.fold_shapes)

    @property
    def buffer_size_time(self) -> int:
        return 2 * self.buffer_size if self.fourier else self.buffer_size

    @property
    def buffer_memory_usage(self) -> float:
        elem = 8 if self.fourier else 4
        return float(np.float32(2 * self.buffer_size * elem) / np.float32(1 << 30))

    @property
    def coord_size(self) -> int:
        return sum(self.ncoords)

    @property
    def coord_memory_usage(self) -> float:
        elem = 8 if self.n_params == 1 else 16
        return float(np.float32(self.coord_size * elem) / np.float32(1 << 30))

    def param_cart_strides(self, lvl: int) -> list[int]:
        c = self.param_counts[lvl]
        s = [1] * len(c)
        for i in range(len(c) - 2, -1, -1):
            s[i] = s[i + 1] * c[i + 1]
        return s

    def compute_param_grid(self, lvl: int) -> list[np.ndarray]:
        return [range_param(lo, hi, d) for (lo, hi), d in zip(self.cfg.param_limits, self.dparams[lvl])]

    @property
    def params_dict(self) -> dict:
        return {n: g for n, g in zip(self.cfg.param_names, self.compute_param_grid(self.n_levels - 1))}

    # -- shift tables -------------------------------------------------------
    def resolve_level_freq(self, lvl: int) -> np.ndarray:
        """taylor_ffa.cpp:249-271 (freq-only): tail unshifted; head shift over
        delta_t = tseg_brute * 2^(l-1); idx = nearest(f_cur, prev grid)."""
        lo, hi = self.cfg.param_limits[0]
        n_cur = self.param_counts[lvl][0]
        n_prev = self.param_counts[lvl - 1][0]
        delta_t = math.ldexp(self.cfg.tseg_brute, lvl - 1)
        f_cur = param_val_at_idx(lo, hi, n_cur, np.arange(n_cur))
        out = np.zeros(n_cur, dtype=FFA_COORD_FREQ_DTYPE)
        out["idx"] = nearest_idx_analytical(f_cur, lo, hi, n_prev).astype(np.uint32)
        out["shift"] = phase_idx(delta_t, f_cur, self.cfg.nbins, 0.0)
        return out

    def _resolve_side(self, lvl: int, latter: int):
        """taylor_ffa.cpp:19-86 (accel) and 88-188 (jerk) for one side.
        nparams 4/5 use the last three params (param_stride offset, :111-121)."""
        cfg = self.cfg
        nb = cfg.nbins
        tseg = math.ldexp(cfg.tseg_brute, lvl - 1)
        delta_t = (float(latter) - 0.5) * tseg
        cur, prev = self.param_counts[lvl], self.param_counts[lvl - 1]
        lims = cfg.param_limits
        if self.n_params == 2:
            (la, lf) = lims
            n_a, n_f = cur
            n_a_p, n_f_p = prev
            half_dt_sq = 0.5 * delta_t * delta_t
            a_cur = param_val_at_idx(la[0], la[1], n_a, np.arange(n_a))[:, None]
            f_cur = param_val_at_idx(lf[0], lf[1], n_f, np.arange(n_f))[None, :]
            v_new = a_cur * delta_t
            d_new = a_cur * half_dt_sq
            idx_a = nearest_idx_analytical(a_cur, la[0], la[1], n_a_p)
            f_new = f_cur * (1.0 - (v_new * KINVCVAL))
            delay = d_new * KINVCVAL
            phase = phase_idx(delta_t, f_cur, nb, delay)
            idx_f = nearest_idx_analytical(f_new, lf[0], lf[1], n_f_p)
            final = idx_a * n_f_p + idx_f
            return final.reshape(-1).astype(np.uint32), phase.reshape(-1)
        if self.n_params in (3, 4, 5):
            po = self.n_params - 3
            for i in range(po):
                if cur[i] != 1 or prev[i] != 1:
                    raise ValueError("crackle/snap grid counts must be 1 for the FFA resolver")
            lj, la, lf = lims[po], lims[po + 1], lims[po + 2]
            n_j, n_a, n_f = cur[po:]
            n_j_p, n_a_p, n_f_p = prev[po:]
            dt_sq = delta_t * delta_t
            dt_cubed = dt_sq * delta_t
            half_dt_sq = 0.5 * dt_sq
            sixth_dt_cubed = dt_cubed / 6.0
            j_cur = param_val_at_idx(lj[0], lj[1], n_j, np.arange(n_j))[:, None, None]
            a_cur = param_val_at_idx(la[0], la[1], n_a, np.arange(n_a))[None, :, None]
            f_cur = param_val_at_idx(lf[0], lf[1], n_f, np.arange(n_f))[None, None, :]
            j_dt = j_cur * delta_t
            half_j_dt_sq = 0.5 * j_cur * dt_sq
            j_sixth = j_cur * sixth_dt_cubed
            idx_j = nearest_idx_analytical(j_cur, lj[0], lj[1], n_j_p)
            a_new = a_cur + j_dt
            v_new = (a_cur * delta_t) + half_j_dt_sq
            d_new = (a_cur * half_dt_sq) + j_sixth
            idx_a = nearest_idx_analytical(a_new, la[0], la[1], n_a_p)
            f_new = f_cur * (1.0 - (v_new * KINVCVAL))
            delay = d_new * KINVCVAL
            phase = phase_idx(delta_t, f_cur, nb, delay)
            idx_f = nearest_idx_analytical(f_new, lf[0], lf[1], n_f_p)
            final = (idx_j * n_a_p * n_f_p + idx_a * n_f_p) + idx_f
            return final.reshape(-1).astype(np.uint32), phase.reshape(-1)
        raise ValueError("unsupported nparams")

    def resolve_level_poly(self, lvl: int) -> np.ndarray:
        """plans.cpp:77-93: tail (delta_t = -tseg_prev/2) and head (+tseg_prev/2)."""
        i_t, s_t = self._resolve_side(lvl, 0)
        i_h, s_h = self._resolve_side(lvl, 1)
        out = np.zeros(i_t.shape[0], dtype=FFA_COORD_DTYPE)
        out["i_tail"], out["shift_tail"], out["i_head"], out["shift_head"] = i_t, s_t, i_h, s_h
        return out

    def resolve_coordinates(self) -> list[np.ndarray]:
        """All levels; level 0 is all-zero (plans.cpp:95-110, 134-150)."""
        res = []
        for lvl in range(self.n_levels):
            if lvl == 0:
                dt = FFA_COORD_FREQ_DTYPE if self.n_params == 1 else FFA_COORD_DTYPE
                res.append(np.zeros(self.ncoords[0], dtype=dt))
            elif self.n_params == 1:
                res.append(self.resolve_level_freq(lvl))
            else:
                res.append(self.resolve_level_poly(lvl))
        return res

    def brute_fold_freqs(self) -> np.ndarray:
        """ffa.cpp:174-177: L0 freq grid (range_param)."""
        return self.compute_param_grid(0)[-1]

    def brute_fold_tref(self) -> float:
        return 0.0 if self.n_params == 1 else self.tsegments[0] / 2.0
