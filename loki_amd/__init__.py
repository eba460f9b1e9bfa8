"""loki_amd: MI355X (gfx950) fast-folding (FFA) hot path for pravirkr/loki.

The package mirrors loki's Python surface for this path:

* ``loki_amd.libloki``   -- ``configs.PulsarSearchConfig``, ``plans.FFAPlan{Base,Time,Fourier}``,
  ``scores.generate_box_width_trials`` (host C++20, src/loki_python.cpp:71-520);
* ``loki_amd.libculoki`` -- ``ffa``, ``fold`` and ``scores`` backed by hand-written gfx950 HIP
  kernels (src/loki_python_cuda.cpp:18-203).

Everything goes through the C-ABI in ``include/loki_amd.h`` (``libloki_amd.so``).  There is no
CPU fallback: importing fails loudly when the native libraries have not been built
(``python loki_amd/build.py``).
"""

from __future__ import annotations

from pathlib import Path

_PKG = Path(__file__).resolve().parent

# torch ships its own HIP runtime under the same SONAME (libamdhip64.so.7) but loads it by an
# unversioned name; if libloki_amd.so were loaded first the process would end up with two HIP
# runtimes and torch could not see the GPU.  Loading torch first makes the loader reuse its
# runtime for libloki_amd.so, so device pointers and hipStream_t handles from torch are valid
# inputs to the C-ABI.
try:
    import torch  # noqa: F401
except ImportError:  # torch is optional: the C-ABI does not need it
    torch = None

try:
    from . import _loki_amd as _native  # noqa: F401
except ImportError as exc:  # pragma: no cover - exercised only on a broken install
    raise ImportError(
        "loki_amd native library is missing or failed to load "
        f"({exc}); build it in-tree with `python loki_amd/build.py`"
    ) from exc

from . import libculoki, libloki  # noqa: E402

LIBRARY_PATH = _PKG / "libloki_amd.so"

__all__ = ["LIBRARY_PATH", "libculoki", "libloki"]
