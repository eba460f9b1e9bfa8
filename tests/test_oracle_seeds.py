"""Pins of the pruning-seed oracle (oracle/seeds.py) on the CPU.

* time linear: an element-by-element loop over the reference's own index formula
  (kernels_cuda.cu:41-48: idx = ibin < shift ? ibin + nbins - shift : ibin - shift) and the
  circular tree index, with the shift rounding edges (s -> u32(s + 0.5), nbins -> 0);
* ascend: the sum over segments of linear seeds on a zero tree, in segment order;
* Fourier: the shift theorem rfft(roll(x, s)) = rfft(x) exp(-2 pi i k s / n) for integer s, so
  the complex seeds of rfft'd rows equal the rfft of the time seeds (pins the phasor's sign and
  the head-only rotation), and a fractional shift against an FP64 DFT of the band-limited
  continuation.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import seeds


def _case(rng, nbins, n_tree=11, n_ffa=7, n_leaves=23):
    tree = rng.standard_normal((n_tree, 2, nbins)).astype(np.float32)
    ffa = rng.standard_normal((n_ffa, 2, nbins)).astype(np.float32)
    itree = rng.integers(0, n_tree, n_leaves).astype(np.uint32)
    iffa = rng.integers(0, n_ffa, n_leaves).astype(np.uint32)
    shift = (rng.random(n_leaves) * nbins).astype(np.float32)
    return tree, ffa, itree, iffa, shift


@pytest.mark.parametrize("nbins", [7, 16, 64])
def test_linear_matches_reference_index_loop(nbins):
    rng = np.random.default_rng(nbins)
    tree, ffa, itree, iffa, shift = _case(rng, nbins)
    shift[:3] = [nbins - 0.4, nbins - 0.6, 0.49]  # -> 0 (nbins), nbins - 1, 0
    n_tree, start = tree.shape[0], 5
    mask = (rng.random(itree.size) > 0.3).astype(np.uint8)
    out0 = np.full((itree.size, 2, nbins), 7.0, np.float32)
    got = seeds.shift_add_linear(tree, itree, mask, ffa, iffa, shift, out0.copy(), nbins, start,
                                 n_tree)
    want = out0.copy()
    for leaf in range(itree.size):
        if not mask[leaf]:
            continue
        s = int(np.float32(shift[leaf]) + np.float32(0.5))
        s = 0 if s == nbins else s
        t = int(itree[leaf]) + start
        t = t - n_tree if t >= n_tree else t
        for q in range(2):
            for b in range(nbins):
                src = b + nbins - s if b < s else b - s
                want[leaf, q, b] = np.float32(tree[t, q, b] + ffa[iffa[leaf], q, src])
    np.testing.assert_array_equal(got, want)
    assert seeds.shift_int(np.float32(nbins - 0.4), nbins) == 0
    assert seeds.shift_int(np.float32(nbins - 0.6), nbins) == nbins - 1


def test_ascend_is_ordered_sum_of_linear_seeds():
    rng = np.random.default_rng(3)
    nbins, nci, nleaves, nsegf, nseg = 32, 9, 17, 5, 4
    ffa = rng.standard_normal((nsegf, nci, 2, nbins)).astype(np.float32)
    iseg = np.array([4, 0, 2, 2], np.uint32)
    iffa = rng.integers(0, nci, (nseg, nleaves)).astype(np.uint32)
    sh = (rng.random((nseg, nleaves)) * nbins).astype(np.float32)
    got = seeds.shift_add_ascend_linear(ffa, iseg, iffa, sh, nbins, nci, nleaves)
    acc = np.zeros((nleaves, 2, nbins), np.float32)
    zero_tree = np.zeros((1, 2, nbins), np.float32)
    for g in range(nseg):
        part = seeds.shift_add_linear(zero_tree, np.zeros(nleaves, np.uint32), None,
                                      ffa[iseg[g]], iffa[g], sh[g],
                                      np.zeros((nleaves, 2, nbins), np.float32), nbins, 0, 1)
        acc = (acc + part).astype(np.float32)
    np.testing.assert_array_equal(got, acc)


@pytest.mark.parametrize("nbins", [16, 33, 64])
def test_fourier_seeds_follow_shift_theorem(nbins):
    rng = np.random.default_rng(10 + nbins)
    tree, ffa, itree, iffa, _ = _case(rng, nbins)
    nbf = nbins // 2 + 1
    shift = rng.integers(0, nbins, itree.size).astype(np.float32)  # integer shifts
    t_time = seeds.shift_add_linear(tree, itree, None, ffa, iffa, shift,
                                    np.zeros((itree.size, 2, nbins), np.float32), nbins, 0,
                                    tree.shape[0])
    tree_c = np.fft.rfft(tree.astype(np.float64), axis=-1)
    ffa_c = np.fft.rfft(ffa.astype(np.float64), axis=-1)
    got = seeds.shift_add_linear_complex(tree_c, itree, None, ffa_c, iffa, shift,
                                         np.zeros((itree.size, 2, nbf), np.complex128), nbf,
                                         nbins, 0, tree.shape[0])
    want = np.fft.rfft(t_time.astype(np.float64), axis=-1)
    np.testing.assert_allclose(got, want, atol=1e-5 * np.abs(want).max())

    # ascend: sum over segments of rfft(roll(row))
    nci, nleaves = ffa.shape[0], 6
    iseg = np.array([0, 0, 0], np.uint32)
    ia = rng.integers(0, nci, (3, nleaves)).astype(np.uint32)
    sa = rng.integers(0, nbins, (3, nleaves)).astype(np.float32)
    got_a = seeds.shift_add_ascend_linear_complex(ffa_c[None], iseg, ia, sa, nbf, nbins, nci,
                                                  nleaves)
    want_t = seeds.shift_add_ascend_linear(ffa[None], iseg, ia, sa, nbins, nci, nleaves)
    np.testing.assert_allclose(got_a, np.fft.rfft(want_t.astype(np.float64), axis=-1),
                               atol=1e-5 * np.abs(got_a).max())


def test_fourier_fractional_shift_delays_a_band_limited_row():
    """A row of harmonics X_k becomes X_k e^{-2 pi i k s/n}: the continuous profile
    x(t) = sum_k Re(c_k X_k e^{2 pi i k t/n}) delayed by s bins."""
    nbins, s = 32, 5.37
    nbf = nbins // 2 + 1
    rng = np.random.default_rng(1)
    x = np.fft.rfft(rng.standard_normal(nbins))
    x[-1] = 0.0  # no Nyquist term: the delayed profile stays real

    def profile(c, t):
        k = np.arange(nbf)
        w = np.where((k == 0), 1.0, 2.0)
        return (w[None, :] * (c[None, :] * np.exp(2j * np.pi * np.outer(t, k) / nbins))).real.sum(1) / nbins

    ffa = np.stack([x, x])[None]  # [1, 2, nbf]
    got = seeds.shift_add_linear_complex(np.zeros_like(ffa), [0], None, ffa, [0],
                                         np.float32([s]), np.zeros((1, 2, nbf), complex), nbf,
                                         nbins, 0, 1)
    t = np.linspace(0, nbins, 97)
    s32 = float(np.float32(s))
    np.testing.assert_allclose(profile(got[0, 0], t), profile(x, t - s32), atol=1e-12)
