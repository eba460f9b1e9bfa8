"""Candidate tables from filtered FFA scores, in the layout of the reference's result writer.

``FFAFreqSweepCUDA::save_results`` (ffa_freq_sweep_cuda.cu:263-340) turns each passing score
index ``trial * nwidths + width`` into a row of ``param_sets``: the final-grid value of every
search parameter (cell centres, psr_utils.hpp:283-293, with the frequency fastest) followed by
the boxcar width; ``FFAResultWriter`` (cands.cpp:385-467) stores the attributes
``ffa_version``, ``param_names`` (the search parameters plus ``width``) and ``scoring_widths``
and the datasets ``snr`` (float32 [n]) and ``param_sets`` (float64 [n, nparams + 1]).

The reference writes HDF5 through HighFive; no HDF5 library is available to this package, so
``write_results`` stores the same attributes and datasets in a NumPy ``.npz`` archive.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

FFA_VERSION = "1.0.0-cpp"


def param_sets(plan, indices, widths) -> np.ndarray:
    """Rows (param_1, ..., param_n, width) for per-width filtered score indices of `plan`."""
    widths = np.asarray(widths)
    nw = widths.size
    idx = np.asarray(indices, dtype=np.int64)
    counts = [int(c) for c in plan.param_counts[-1]]
    limits = np.asarray(plan.config.param_limits, dtype=np.float64)
    coord = idx // nw
    if coord.size and coord.max() >= int(np.prod(counts)):
        raise ValueError("score index beyond the final grid")
    cols = np.unravel_index(coord, counts)
    out = np.empty((idx.size, len(counts) + 1), dtype=np.float64)
    for j, (c, (lo, hi), col) in enumerate(zip(counts, limits, cols)):
        out[:, j] = _cell_centre(float(lo), float(hi), c, col)
    out[:, -1] = widths[idx % nw].astype(np.float64)
    return out


def _cell_centre(vmin: float, vmax: float, count: int, i) -> np.ndarray:
    """psr_utils.hpp:283-293 get_param_val_at_idx: vmin + (vmax - vmin) / count * (i + 0.5),
    the midpoint of the range for a single cell."""
    i = np.asarray(i, dtype=np.float64)
    if count <= 1:
        return np.full(i.shape, (vmax + vmin) / 2.0)
    return vmin + ((vmax - vmin) / float(count)) * (i + 0.5)


def write_results(path, param_names, scoring_widths, snr, param_sets_arr) -> Path:
    """Writes one result archive; returns its path (a ``.npz`` suffix is added if missing)."""
    path = Path(path)
    if path.suffix != ".npz":
        path = path.with_suffix(path.suffix + ".npz")
    snr = np.asarray(snr, dtype=np.float32)
    ps = np.asarray(param_sets_arr, dtype=np.float64).reshape(snr.size, -1)
    names = list(param_names)
    if names[-1] != "width":
        names = names + ["width"]
    if ps.shape[1] != len(names):
        raise ValueError("param_sets size does not match the expected dimension")
    np.savez(path, ffa_version=np.array(FFA_VERSION), param_names=np.array(names),
             scoring_widths=np.asarray(scoring_widths, dtype=np.uint64), snr=snr, param_sets=ps)
    return path


def read_results(path) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


__all__ = ["FFA_VERSION", "param_sets", "read_results", "write_results"]
