"""Oracle restatement of the pruning seeds built from final-level FFA folds.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows kernels.cpp:1328-1452 (shift_add_linear_batch, shift_add_linear_complex_batch,
shift_add_ascend_linear_batch, shift_add_ascend_linear_complex_batch) and the per-row helpers
kernels.cpp:118-176 (right rotation of the head by u32(shift + 0.5f), nbins -> 0, then an FP32
add) and kernels.cpp:369-520 (head times exp(-2 pi i k shift / nbins)); the validation mask and
the circular tree index follow the CUDA kernels (kernels_cuda.cu:16-65, 141-203).

The time-domain seeds are numpy float32 element-wise adds in the reference's order (0 + a_0 +
a_1 + ...), hence bit-exact definitions.  The Fourier seeds are evaluated in FP64 (the
reference's own rounding comes from an xsimd recurrence on the CPU and __sincosf on the GPU),
so they are compared within a tolerance.

Pinned by tests/test_oracle_seeds.py: element-by-element loops over the reference's index
formula, the shift rounding edge cases, ascend = sum of linear seeds, and the Fourier shift
theorem rfft(roll(x, s)) = rfft(x) exp(-2 pi i k s / n) for integer s.
"""

from __future__ import annotations

import numpy as np


def shift_int(phase_shift, nbins: int) -> np.ndarray:
    """kernels.cpp:124-127: SizeType(s + 0.5f) with nbins -> 0 (the add is in float32)."""
    s = (np.asarray(phase_shift, np.float32) + np.float32(0.5)).astype(np.uint64)
    return np.where(s == nbins, 0, s).astype(np.int64)


def _rotate_rows(rows: np.ndarray, shifts: np.ndarray) -> np.ndarray:
    """rows [n, 2, nbins]: out[i, :, j] = rows[i, :, (j - shift_i) mod nbins]."""
    nbins = rows.shape[-1]
    j = np.arange(nbins)
    src = (j[None, :] - shifts[:, None]) % nbins
    return np.take_along_axis(rows, np.broadcast_to(src[:, None, :], rows.shape), axis=2)


def tree_physical(indices_tree, start: int, capacity: int) -> np.ndarray:
    """kernels_cuda.cu:46-49: uint32 (index + start), less capacity once."""
    t = (np.asarray(indices_tree, np.uint64) + np.uint64(start)) % np.uint64(1 << 32)
    return np.where(t >= capacity, t - np.uint64(capacity), t).astype(np.int64)


def shift_add_linear(folds_tree, indices_tree, mask, folds_ffa, indices_ffa, phase_shift,
                     out, nbins: int, start: int, capacity: int) -> np.ndarray:
    """kernels.cpp:1328-1355 + kernels_cuda.cu:16-65; `out` [n_leaves, 2, nbins] is updated
    in place for valid leaves and returned."""
    tree = np.asarray(folds_tree, np.float32).reshape(-1, 2, nbins)
    ffa = np.asarray(folds_ffa, np.float32).reshape(-1, 2, nbins)
    n = len(indices_ffa)
    valid = np.ones(n, bool) if mask is None else np.asarray(mask) != 0
    t = tree[tree_physical(indices_tree, start, capacity)]
    h = _rotate_rows(ffa[np.asarray(indices_ffa, np.int64)], shift_int(phase_shift, nbins))
    res = (t + h).astype(np.float32)
    out = np.asarray(out, np.float32).reshape(n, 2, nbins)
    out[valid] = res[valid]
    return out


def shift_add_ascend_linear(folds_ffa, indices_segment, indices_ffa, phase_shift, nbins: int,
                            n_coords_init: int, n_leaves: int) -> np.ndarray:
    """kernels.cpp:1390-1416: per leaf, zero then += each segment's rotated row in order."""
    ffa = np.asarray(folds_ffa, np.float32).reshape(-1, n_coords_init, 2, nbins)
    iffa = np.asarray(indices_ffa, np.int64).reshape(-1, n_leaves)
    sh = np.asarray(phase_shift, np.float32).reshape(-1, n_leaves)
    acc = np.zeros((n_leaves, 2, nbins), np.float32)
    for g, seg in enumerate(np.asarray(indices_segment, np.int64)):
        acc += _rotate_rows(ffa[seg][iffa[g]], shift_int(sh[g], nbins))
    return acc


def _phasor(phase_shift, nbf: int, nbins: int) -> np.ndarray:
    """exp(-2 pi i k s / nbins), k = 0..nbf-1, in FP64: [n, nbf]."""
    k = np.arange(nbf, dtype=np.float64)
    s = np.asarray(phase_shift, np.float32).astype(np.float64)
    return np.exp(-2j * np.pi * (s[:, None] * k[None, :]) / nbins)


def shift_add_linear_complex(folds_tree, indices_tree, mask, folds_ffa, indices_ffa,
                             phase_shift, out, nbins_f: int, nbins: int, start: int,
                             capacity: int) -> np.ndarray:
    """kernels.cpp:1357-1388 + kernels_cuda.cu:141-203, in FP64; returns complex128
    [n_leaves, 2, nbins_f] (invalid leaves keep `out`)."""
    tree = np.asarray(folds_tree).reshape(-1, 2, nbins_f).astype(np.complex128)
    ffa = np.asarray(folds_ffa).reshape(-1, 2, nbins_f).astype(np.complex128)
    n = len(indices_ffa)
    valid = np.ones(n, bool) if mask is None else np.asarray(mask) != 0
    w = _phasor(phase_shift, nbins_f, nbins)
    res = tree[tree_physical(indices_tree, start, capacity)] + \
        ffa[np.asarray(indices_ffa, np.int64)] * w[:, None, :]
    out = np.asarray(out).reshape(n, 2, nbins_f).astype(np.complex128)
    out[valid] = res[valid]
    return out


def shift_add_ascend_linear_complex(folds_ffa, indices_segment, indices_ffa, phase_shift,
                                    nbins_f: int, nbins: int, n_coords_init: int,
                                    n_leaves: int) -> np.ndarray:
    """kernels.cpp:1418-1452 in FP64."""
    ffa = np.asarray(folds_ffa).reshape(-1, n_coords_init, 2, nbins_f).astype(np.complex128)
    iffa = np.asarray(indices_ffa, np.int64).reshape(-1, n_leaves)
    sh = np.asarray(phase_shift, np.float32).reshape(-1, n_leaves)
    acc = np.zeros((n_leaves, 2, nbins_f), np.complex128)
    for g, seg in enumerate(np.asarray(indices_segment, np.int64)):
        acc += ffa[seg][iffa[g]] * _phasor(sh[g], nbins_f, nbins)[:, None, :]
    return acc
