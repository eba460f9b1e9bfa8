"""Pruning seeds on the GPU through the C-ABI (kernels_cuda.cu:16-277 shift_add_*_batch_cuda).

Time-domain seeds are compared bit-for-bit with the oracle (oracle/seeds.py, the reference's
CPU order of FP32 adds); Fourier seeds within 2e-6 of the FP64 oracle relative to the largest
input magnitude times the number of terms.  Inputs are random rows and, in one case, the final
level of a real FFA run.  Covered: validation masks (masked leaves keep their output), the
circular tree index with its wrap, shift rounding edges (nbins - 0.4 -> 0), non-power-of-two
nbins, no leaves, and out-of-range indices (NaN output, no out-of-bounds read).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_vp, _sz, _u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32


def _lib():
    import loki_amd

    lib = ctypes.CDLL(str(loki_amd.LIBRARY_PATH))
    lib.loki_last_error.restype = ctypes.c_char_p
    lib.loki_shift_add_linear_batch_device.argtypes = [_vp] * 7 + [_sz, _sz, _u32, _u32, _sz, _vp]
    lib.loki_shift_add_linear_complex_batch_device.argtypes = [_vp] * 7 + [_sz, _sz, _sz, _u32,
                                                                            _u32, _sz, _vp]
    lib.loki_shift_add_ascend_linear_batch_device.argtypes = [_vp] * 5 + [_sz] * 5 + [_vp]
    lib.loki_shift_add_ascend_linear_complex_batch_device.argtypes = [_vp] * 5 + [_sz] * 6 + [_vp]
    return lib


def _dev(a):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:  # same bits; torch's uint32 tensors have few operations
        a = a.view(np.int32)
    return torch.from_numpy(a).to("cuda:0")


def _ptr(t):
    return None if t is None else t.data_ptr()


def _ok(lib, rc):
    assert rc == 0, lib.loki_last_error()


def _c64(a):
    return np.ascontiguousarray(a.astype(np.complex64)).view(np.float32)


def _linear_inputs(rng, nbins, n_tree, n_ffa, n_leaves, start):
    tree = rng.standard_normal((n_tree, 2, nbins)).astype(np.float32)
    ffa = rng.standard_normal((n_ffa, 2, nbins)).astype(np.float32) * 3
    itree = rng.integers(0, n_tree, n_leaves).astype(np.uint32)
    # logical indices: physical = (logical + start) mod n_tree, so stored logical = phys - start
    itree = ((itree.astype(np.int64) - start) % n_tree).astype(np.uint32)
    iffa = rng.integers(0, n_ffa, n_leaves).astype(np.uint32)
    shift = (rng.random(n_leaves) * nbins).astype(np.float32)
    shift[:2] = [nbins - 0.4, 0.5]
    mask = (rng.random(n_leaves) > 0.25).astype(np.uint8)
    return tree, ffa, itree, iffa, shift, mask


@pytest.mark.parametrize("nbins,n_leaves", [(64, 5000), (100, 777), (32, 1 << 16), (8, 3),
                                             (33, 1000), (30, 513)])
def test_linear_time_bit_exact(nbins, n_leaves):
    import torch

    from oracle import seeds

    lib = _lib()
    rng = np.random.default_rng(nbins + n_leaves)
    n_tree, n_ffa, start = 257, 131, 200
    tree, ffa, itree, iffa, shift, mask = _linear_inputs(rng, nbins, n_tree, n_ffa, n_leaves,
                                                         start)
    out0 = np.full((n_leaves, 2, nbins), -3.0, np.float32)
    d = [_dev(x) for x in (tree, itree, mask, ffa, iffa, shift, out0)]
    _ok(lib, lib.loki_shift_add_linear_batch_device(*map(_ptr, d), nbins, n_leaves, start,
                                                    n_tree, n_ffa, None))
    torch.cuda.synchronize()
    want = seeds.shift_add_linear(tree, itree, mask, ffa, iffa, shift, out0.copy(), nbins, start,
                                  n_tree)
    np.testing.assert_array_equal(d[6].cpu().numpy(), want)
    # no mask: every leaf
    _ok(lib, lib.loki_shift_add_linear_batch_device(_ptr(d[0]), _ptr(d[1]), None, *map(_ptr, d[3:]),
                                                    nbins, n_leaves, start, n_tree, n_ffa, None))
    torch.cuda.synchronize()
    want = seeds.shift_add_linear(tree, itree, None, ffa, iffa, shift, out0.copy(), nbins, start,
                                  n_tree)
    np.testing.assert_array_equal(d[6].cpu().numpy(), want)


@pytest.mark.parametrize("nbins,nleaves,nseg", [(64, 3000, 4), (48, 513, 7), (32, 1 << 15, 2),
                                               (30, 700, 5), (64, 100, 1)])
def test_ascend_time_bit_exact(nbins, nleaves, nseg):
    import torch

    from oracle import seeds

    lib = _lib()
    rng = np.random.default_rng(nbins * nseg)
    nci, nsegf = 97, nseg + 3
    ffa = rng.standard_normal((nsegf, nci, 2, nbins)).astype(np.float32)
    iseg = rng.permutation(nsegf)[:nseg].astype(np.uint32)
    iffa = rng.integers(0, nci, (nseg, nleaves)).astype(np.uint32)
    sh = (rng.random((nseg, nleaves)) * nbins).astype(np.float32)
    out = np.full((nleaves, 2, nbins), np.nan, np.float32)  # fully overwritten
    d = [_dev(x) for x in (ffa, iseg, iffa, sh, out)]
    _ok(lib, lib.loki_shift_add_ascend_linear_batch_device(*map(_ptr, d), nbins, nci, nleaves,
                                                           nseg, nsegf, None))
    torch.cuda.synchronize()
    want = seeds.shift_add_ascend_linear(ffa, iseg, iffa, sh, nbins, nci, nleaves)
    np.testing.assert_array_equal(d[4].cpu().numpy(), want)


def test_linear_time_unaligned_buffers_take_the_scalar_path():
    import torch

    from oracle import seeds

    lib = _lib()
    rng = np.random.default_rng(44)
    nbins, n_tree, n_ffa, n, start = 64, 50, 40, 300, 7
    tree, ffa, itree, iffa, shift, mask = _linear_inputs(rng, nbins, n_tree, n_ffa, n, start)
    # each array one float past a 16-byte boundary
    buf_t = _dev(np.concatenate([[0.0], tree.ravel()]).astype(np.float32))
    buf_f = _dev(np.concatenate([[0.0], ffa.ravel()]).astype(np.float32))
    buf_o = _dev(np.zeros(1 + n * 2 * nbins, np.float32))
    di, dm, dfi, ds = _dev(itree), _dev(mask), _dev(iffa), _dev(shift)
    _ok(lib, lib.loki_shift_add_linear_batch_device(
        buf_t.data_ptr() + 4, di.data_ptr(), dm.data_ptr(), buf_f.data_ptr() + 4, dfi.data_ptr(),
        ds.data_ptr(), buf_o.data_ptr() + 4, nbins, n, start, n_tree, n_ffa, None))
    torch.cuda.synchronize()
    want = seeds.shift_add_linear(tree, itree, mask, ffa, iffa, shift,
                                  np.zeros((n, 2, nbins), np.float32), nbins, start, n_tree)
    np.testing.assert_array_equal(buf_o.cpu().numpy()[1:].reshape(n, 2, nbins), want)


def test_seeds_from_a_real_ffa_fold():
    """folds_ffa = the final level of an accelerated search, as the pruning stage reads it."""
    import torch

    from loki_amd import libculoki, libloki
    from oracle import seeds

    lib = _lib()
    cfg = libloki.configs.PulsarSearchConfig(
        nsamps=2**16, tsamp=1e-4, nbins=64, eta=1.0, bseg_ffa=2**14,
        param_limits=np.array([[-20.0, 20.0], [10.0, 30.0]]), use_fourier=False)
    rng = np.random.default_rng(8)
    ts_e = rng.standard_normal(2**16).astype(np.float32)
    ts_v = np.ones(2**16, np.float32)
    fold, plan = libculoki.ffa.compute_ffa_time_cuda(ts_e, ts_v, cfg)
    nseg_f, nci = int(plan.nsegments[-1]), int(plan.ncoords[-1])
    assert nseg_f == 4
    fold = fold.reshape(nseg_f, nci, 2, 64)
    nleaves, nseg = 4096, nseg_f
    iseg = np.arange(nseg, dtype=np.uint32)
    iffa = rng.integers(0, nci, (nseg, nleaves)).astype(np.uint32)
    sh = (rng.random((nseg, nleaves)) * 64).astype(np.float32)
    d = [_dev(x) for x in (fold, iseg, iffa, sh, np.zeros((nleaves, 2, 64), np.float32))]
    _ok(lib, lib.loki_shift_add_ascend_linear_batch_device(*map(_ptr, d), 64, nci, nleaves,
                                                           nseg, nseg_f, None))
    torch.cuda.synchronize()
    want = seeds.shift_add_ascend_linear(fold, iseg, iffa, sh, 64, nci, nleaves)
    np.testing.assert_array_equal(d[4].cpu().numpy(), want)


@pytest.mark.parametrize("nbins,n_leaves", [(64, 4000), (100, 333), (33, 64)])
def test_linear_complex_within_tolerance(nbins, n_leaves):
    import torch

    from oracle import seeds

    lib = _lib()
    rng = np.random.default_rng(nbins)
    nbf = nbins // 2 + 1
    n_tree, n_ffa, start = 129, 71, 100
    tree = (rng.standard_normal((n_tree, 2, nbf)) + 1j * rng.standard_normal((n_tree, 2, nbf)))
    ffa = (rng.standard_normal((n_ffa, 2, nbf)) + 1j * rng.standard_normal((n_ffa, 2, nbf))) * 20
    tree, ffa = tree.astype(np.complex64), ffa.astype(np.complex64)
    itree = ((rng.integers(0, n_tree, n_leaves) - start) % n_tree).astype(np.uint32)
    iffa = rng.integers(0, n_ffa, n_leaves).astype(np.uint32)
    shift = (rng.random(n_leaves) * nbins).astype(np.float32)
    shift[0] = nbins - 1e-3
    mask = (rng.random(n_leaves) > 0.2).astype(np.uint8)
    out0 = np.full((n_leaves, 2, nbf), 5 + 5j, np.complex64)
    d = [_dev(_c64(tree)), _dev(itree), _dev(mask), _dev(_c64(ffa)), _dev(iffa), _dev(shift),
         _dev(_c64(out0))]
    _ok(lib, lib.loki_shift_add_linear_complex_batch_device(*map(_ptr, d), nbf, nbins, n_leaves,
                                                            start, n_tree, n_ffa, None))
    torch.cuda.synchronize()
    got = d[6].cpu().numpy().view(np.complex64).reshape(n_leaves, 2, nbf)
    want = seeds.shift_add_linear_complex(tree, itree, mask, ffa, iffa, shift, out0, nbf, nbins,
                                          start, n_tree)
    scale = np.abs(ffa).max() + np.abs(tree).max()
    err = np.abs(got - want).max()
    print(f"nbins {nbins}: max |err| / scale = {err / scale:.2e}")
    assert err <= 2e-6 * scale
    np.testing.assert_array_equal(got[mask == 0], out0[mask == 0])


@pytest.mark.parametrize("nbins,nleaves,nseg", [(64, 2000, 4), (32, 300, 9)])
def test_ascend_complex_within_tolerance(nbins, nleaves, nseg):
    import torch

    from oracle import seeds

    lib = _lib()
    rng = np.random.default_rng(nbins + nseg)
    nbf = nbins // 2 + 1
    nci, nsegf = 53, nseg + 1
    ffa = (rng.standard_normal((nsegf, nci, 2, nbf)) +
           1j * rng.standard_normal((nsegf, nci, 2, nbf))).astype(np.complex64) * 10
    iseg = rng.permutation(nsegf)[:nseg].astype(np.uint32)
    iffa = rng.integers(0, nci, (nseg, nleaves)).astype(np.uint32)
    sh = (rng.random((nseg, nleaves)) * nbins).astype(np.float32)
    d = [_dev(_c64(ffa)), _dev(iseg), _dev(iffa), _dev(sh),
         _dev(np.zeros(nleaves * 2 * nbf * 2, np.float32))]
    _ok(lib, lib.loki_shift_add_ascend_linear_complex_batch_device(
        *map(_ptr, d), nbf, nbins, nci, nleaves, nseg, nsegf, None))
    torch.cuda.synchronize()
    got = d[4].cpu().numpy().view(np.complex64).reshape(nleaves, 2, nbf)
    want = seeds.shift_add_ascend_linear_complex(ffa, iseg, iffa, sh, nbf, nbins, nci, nleaves)
    scale = np.abs(ffa).max() * nseg
    err = np.abs(got - want).max()
    print(f"nbins {nbins} nseg {nseg}: max |err| / scale = {err / scale:.2e}")
    assert err <= 2e-6 * scale


def test_out_of_range_indices_give_nan_and_empty_is_noop():
    import torch

    lib = _lib()
    nbins, n_tree, n_ffa, n = 16, 4, 3, 5
    tree = np.ones((n_tree, 2, nbins), np.float32)
    ffa = np.ones((n_ffa, 2, nbins), np.float32)
    itree = np.array([0, 1, 9, 2, 3], np.uint32)   # 9 + 0 >= 4 even after one wrap
    iffa = np.array([0, 3, 0, 2, 1], np.uint32)    # 3 == n_ffa
    shift = np.zeros(n, np.float32)
    d = [_dev(x) for x in (tree, itree, ffa, iffa, shift, np.zeros((n, 2, nbins), np.float32))]
    _ok(lib, lib.loki_shift_add_linear_batch_device(_ptr(d[0]), _ptr(d[1]), None, _ptr(d[2]),
                                                    _ptr(d[3]), _ptr(d[4]), _ptr(d[5]), nbins, n,
                                                    0, n_tree, n_ffa, None))
    torch.cuda.synchronize()
    got = d[5].cpu().numpy()
    assert np.isnan(got[1]).all() and np.isnan(got[2]).all()
    np.testing.assert_array_equal(got[[0, 3, 4]], 2.0)
    # ascend: a segment index beyond the fold's segments
    iseg = _dev(np.array([0, 5], np.uint32))
    ia = _dev(np.zeros((2, n), np.uint32))
    sa = _dev(np.zeros((2, n), np.float32))
    out = _dev(np.zeros((n, 2, nbins), np.float32))
    _ok(lib, lib.loki_shift_add_ascend_linear_batch_device(_ptr(d[2]), _ptr(iseg), _ptr(ia),
                                                           _ptr(sa), _ptr(out), nbins, n_ffa, n,
                                                           2, 1, None))
    torch.cuda.synchronize()
    assert np.isnan(out.cpu().numpy()).all()
    # zero leaves: nothing launched, success
    _ok(lib, lib.loki_shift_add_linear_batch_device(None, None, None, None, None, None, None,
                                                    nbins, 0, 0, 1, 1, None))
    assert lib.loki_shift_add_linear_complex_batch_device(
        _ptr(d[0]), _ptr(d[1]), None, _ptr(d[2]), _ptr(d[3]), _ptr(d[4]), _ptr(d[5]), 8, nbins,
        1, 0, n_tree, n_ffa, None) != 0  # nbins_f must be nbins/2 + 1
