"""``loki.libculoki``-compatible GPU modules (ffa, fold, scores) on gfx950.

Mirrors src/loki_python_cuda.cpp:18-203 and src/loki_templates_cuda.cuh:21-51: the same function
and class names, argument names, defaults and return types.  ``FFACUDA`` keeps the reference's
name so that callers switch backends by changing the import only.
"""

from __future__ import annotations

import sys
import types

from . import _loki_amd as _n

ffa = _n.ffa
fold = _n.fold
scores = types.ModuleType(__name__ + ".scores", "Scores submodule")
scores.snr_boxcar_3d = _n.scores.snr_boxcar_3d
scores.snr_boxcar_3d_max = _n.scores.snr_boxcar_3d_max
scores.snr_boxcar_2d_max = _n.scores.snr_boxcar_2d_max
# score_and_filter[_max]_cuda_d (score_cuda.cu:382-443): (indices, snr) above a threshold
scores.snr_boxcar_3d_filter = _n.scores.snr_boxcar_3d_filter

# importable by dotted name (``from loki_amd.libculoki.ffa import FFACUDA``)
for _name, _mod in (("ffa", ffa), ("fold", fold), ("scores", scores)):
    sys.modules[f"{__name__}.{_name}"] = _mod
del _name, _mod

__all__ = ["ffa", "fold", "scores"]
