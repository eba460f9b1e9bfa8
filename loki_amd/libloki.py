"""``loki.libloki``-compatible host modules (configs, plans, scores, ffa).

Mirrors src/loki_python.cpp:71-78 and 275-592.  ``configs`` and ``plans`` are implemented in
C++20 behind the C-ABI of ``include/loki_amd.h``.  ``ffa`` keeps the CPU module's names and
signatures (src/loki_python.cpp:518-592, src/loki_templates.hpp:47-74) so that a caller such as
``loki.search.ffa_search`` (src/loki/search.py:7,41-47) switches backends by its import alone;
the work runs on the gfx950 HIP path of device 0.  ``show_progress`` is accepted and ignored
(no progress bars), as is ``quiet`` (the HIP path does not log).
"""

from __future__ import annotations

import sys
import types

from . import _loki_amd as _n

configs = _n.configs
plans = _n.plans
scores = types.ModuleType(__name__ + ".scores", "Scores submodule")
scores.generate_box_width_trials = _n.scores.generate_box_width_trials


class FFATime:
    """``libloki.ffa.FFATime(cfg, show_progress=True)`` (loki_templates.hpp:47-61):
    ``.plan`` and ``.execute(ts_e, ts_v, fold)`` with ``fold`` a float32 array of exactly
    ``plan.buffer_size`` elements (ffa.cpp:83-85); the final level is left in its first
    ``plan.fold_size`` elements."""

    def __init__(self, cfg, show_progress: bool = True):
        del show_progress
        self._ffa = _n.ffa.FFACUDA(cfg, device_id=0)

    @property
    def plan(self):
        return self._ffa.plan

    def execute(self, ts_e, ts_v, fold):
        self._ffa.execute(ts_e, ts_v, fold)


class FFAFourier:
    """``libloki.ffa.FFAFourier(cfg, show_progress=True)`` (loki_templates.hpp:47-74):
    ``.execute(ts_e, ts_v, fold)`` with a complex64 fold of ``plan.buffer_size`` elements
    (Fourier folds), or a float32 fold of ``2 * plan.buffer_size`` elements (the final level
    returned to the time domain)."""

    def __init__(self, cfg, show_progress: bool = True):
        del show_progress
        self._ffa = _n.ffa.FFACOMPLEXCUDA(cfg, device_id=0)

    @property
    def plan(self):
        return self._ffa.plan

    def execute(self, ts_e, ts_v, fold):
        self._ffa.execute(ts_e, ts_v, fold)


def compute_ffa_time(ts_e, ts_v, cfg, quiet: bool = False, show_progress: bool = False):
    """(fold [fold_size] float32, FFAPlanTime) -- loki_python.cpp:523-535."""
    del show_progress
    return _n.ffa.compute_ffa_time_cuda(ts_e, ts_v, cfg, device_id=0, quiet=quiet)


def compute_ffa_fourier(ts_e, ts_v, cfg, quiet: bool = False, show_progress: bool = False):
    """(fold [fold_size] complex64, FFAPlanFourier) -- loki_python.cpp:537-549."""
    del show_progress
    return _n.ffa.compute_ffa_fourier_cuda(ts_e, ts_v, cfg, device_id=0, quiet=quiet)


def compute_ffa_fourier_return_to_time(ts_e, ts_v, cfg, quiet: bool = False,
                                       show_progress: bool = False):
    """(fold [2 * fold_size] float32, FFAPlanTime) -- loki_python.cpp:551-564."""
    del show_progress
    return _n.ffa.compute_ffa_fourier_return_to_time_cuda(ts_e, ts_v, cfg, device_id=0,
                                                         quiet=quiet)


def compute_ffa_scores(ts_e, ts_v, cfg, quiet: bool = False, show_progress: bool = False):
    """(scores [ncoords_final * len(cfg.score_widths)] float32, FFAPlanTime) --
    loki_python.cpp:566-578; the time FFA, or the Fourier FFA returned to the time domain when
    ``cfg.use_fourier``, scored at ``cfg.score_widths`` (ffa.cpp:444-466)."""
    del show_progress
    return _n.ffa.compute_ffa_scores_cuda(ts_e, ts_v, cfg, device_id=0, quiet=quiet)


class FFAFreqSweep:
    """``libloki.ffa.FFAFreqSweep(cfg, show_progress=True)`` (loki_python.cpp:580-592):
    ``.execute(ts_e, ts_v, outdir, file_prefix="test")`` plans the frequency regions, runs
    each chunk's FFA + score-and-filter at ``cfg.snr_min`` and writes the candidate archive."""

    def __init__(self, cfg, show_progress: bool = True):
        del show_progress
        self._sweep = _n.ffa.FFAFreqSweepCUDA(cfg, device_id=0)

    def execute(self, ts_e, ts_v, outdir, file_prefix: str = "test"):
        return self._sweep.execute(ts_e, ts_v, outdir, file_prefix)


ffa = types.ModuleType(__name__ + ".ffa", "FFA submodule")
for _obj in (FFATime, FFAFourier, FFAFreqSweep, compute_ffa_time, compute_ffa_fourier,
             compute_ffa_fourier_return_to_time, compute_ffa_scores):
    setattr(ffa, _obj.__name__, _obj)

# ``from loki_amd.libloki.ffa import compute_ffa_scores`` (the form search.py:7 uses) imports
# the submodules by dotted name, so they are registered as modules.
for _name, _mod in (("configs", configs), ("plans", plans), ("scores", scores), ("ffa", ffa)):
    sys.modules[f"{__name__}.{_name}"] = _mod
del _obj, _name, _mod

__all__ = ["configs", "ffa", "plans", "scores"]
