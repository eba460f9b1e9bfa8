"""The allocation-free brute fold (loki_brute_fold_time_device_noalloc) against the allocating
one (loki_brute_fold_time_device), through the C-ABI on cuda:0.

Bar: bit-identical folds.  Two calls in a row, the second with a smaller input (fewer samples,
frequencies and bins), so stale contents of the shared device scratch from the first call
would show.  Also: the run path (segments longer than 4096 samples, whose scratch holds only
the frequencies), the error for tables beyond the 64 MiB scratch, and that the list path and
the oracle agree on the first case.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    import loki_amd

    lib = ctypes.CDLL(str(loki_amd.LIBRARY_PATH))
    args = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
            ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
            ctypes.c_void_p]
    for fn in (lib.loki_brute_fold_time_device, lib.loki_brute_fold_time_device_noalloc):
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.loki_last_error.restype = ctypes.c_char_p
    return lib


def _fold_both(lib, nsamps, seglen, nbins, freqs, tsamp, t_ref, seed):
    import torch

    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    te = torch.from_numpy(rng.standard_normal(nsamps, dtype=np.float32)).to(dev)
    tv = torch.from_numpy((1.0 + 0.25 * rng.random(nsamps)).astype(np.float32)).to(dev)
    n_out = (nsamps // seglen) * freqs.size * 2 * nbins
    outs = []
    for fn in (lib.loki_brute_fold_time_device, lib.loki_brute_fold_time_device_noalloc):
        out = torch.full((n_out,), float("nan"), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        rc = fn(te.data_ptr(), tv.data_ptr(), nsamps, freqs.ctypes.data, freqs.size, seglen,
                nbins, tsamp, t_ref, out.data_ptr(), None)
        assert rc == 0, lib.loki_last_error().decode()
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    return te.cpu().numpy(), tv.cpu().numpy(), outs[0], outs[1]


def test_no_alloc_brute_fold_equals_allocating_one_twice(oracle):
    lib = _lib()
    tsamp = 64e-6
    # first call: 2^18 samples, 1024-sample segments, 600 frequencies, 64 bins
    freqs1 = np.linspace(40.0, 55.0, 600)
    e1, v1, want1, got1 = _fold_both(lib, 2**18, 1024, 64, freqs1, tsamp, 0.0, 1)
    np.testing.assert_array_equal(got1, want1)
    ref = oracle.brute_fold_time(e1, v1, freqs1, 1024, 64, tsamp, 0.0)
    np.testing.assert_array_equal(got1, ref)
    # second call, smaller: 2^16 samples, 512-sample segments, 77 frequencies, 16 bins
    freqs2 = np.linspace(100.0, 104.0, 77)
    _, _, want2, got2 = _fold_both(lib, 2**16, 512, 16, freqs2, tsamp, 0.016, 2)
    np.testing.assert_array_equal(got2, want2)
    assert np.isfinite(got2).all()


def test_no_alloc_brute_fold_run_path_and_limit():
    lib = _lib()
    tsamp = 64e-6
    # run path: 8192-sample segments
    freqs = np.linspace(5.0, 6.0, 40)
    _, _, want, got = _fold_both(lib, 2**16, 8192, 32, freqs, tsamp, 0.0, 3)
    np.testing.assert_array_equal(got, want)
    # bucket tables beyond the 64 MiB scratch: a clear error, nothing written
    import torch

    nf, seglen, nbins = 20000, 4096, 1024
    big = np.linspace(1.0, 2.0, nf)
    dev = torch.device("cuda:0")
    te = torch.zeros(seglen, dtype=torch.float32, device=dev)
    out = torch.zeros(1, dtype=torch.float32, device=dev)
    rc = lib.loki_brute_fold_time_device_noalloc(te.data_ptr(), te.data_ptr(), seglen,
                                                 big.ctypes.data, nf, seglen, nbins, tsamp, 0.0,
                                                 out.data_ptr(), None)
    assert rc != 0
    assert "scratch" in lib.loki_last_error().decode()
