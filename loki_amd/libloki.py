"""``loki.libloki``-compatible host modules (configs, plans, scores.generate_box_width_trials).

Mirrors src/loki_python.cpp:71-78 and 275-520; the classes are implemented in C++20 behind the
C-ABI of ``include/loki_amd.h``.
"""

from __future__ import annotations

from types import SimpleNamespace

from . import _loki_amd as _n

configs = _n.configs
plans = _n.plans
scores = SimpleNamespace(generate_box_width_trials=_n.scores.generate_box_width_trials)

__all__ = ["configs", "plans", "scores"]
