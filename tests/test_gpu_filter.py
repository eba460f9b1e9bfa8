"""Score-and-filter on the GPU (score_cuda.cu:382-443 score_and_filter[_max]_cuda_d).

The filtered outputs are checked against the full scores of the same path, which the parity
tests hold bit-exact to the oracle (time domain) or within the Fourier bound: the filter must
return exactly the (index, S/N) pairs with S/N > threshold, the S/N bits unchanged, for the
standalone scorer, the fused time-domain last level (FP32 and reduced storage), the unfused
path and the fused Fourier path.  Over-full outputs report the true count and write only
`capacity` entries, all of them valid candidates.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from test_gpu_parity import _cfgs, _series

pytestmark = pytest.mark.gpu


def _expected(full, nw, threshold, find_max):
    full = np.asarray(full, np.float32).reshape(-1, nw)
    if find_max:
        best = full.max(axis=1)
        idx = np.nonzero(best > np.float32(threshold))[0]
        return idx.astype(np.uint32), best[idx]
    flat = full.reshape(-1)
    idx = np.nonzero(flat > np.float32(threshold))[0]
    return idx.astype(np.uint32), flat[idx]


def _check(got, want, label):
    gi, gs = got
    wi, ws = want
    assert gi.dtype == np.uint32 and gs.dtype == np.float32
    print(f"{label}: {gi.size} candidates")
    np.testing.assert_array_equal(gi, wi)
    np.testing.assert_array_equal(gs, ws)


@pytest.mark.parametrize("nbins", [8, 32, 64, 128, 48, 100])
@pytest.mark.parametrize("find_max", [False, True])
def test_standalone_filter_equals_thresholded_scores(nbins, find_max):
    from loki_amd import libculoki

    rng = np.random.default_rng(nbins)
    nprof = 3001
    e = rng.standard_normal((nprof, 1, nbins)) * 2.0
    e[::7, 0, : max(1, nbins // 8)] += 6.0  # some strong pulses
    v = np.full((nprof, 1, nbins), 4.0) + rng.random((nprof, 1, nbins))
    folds = np.concatenate([e, v], axis=1).astype(np.float32)
    widths = np.array([w for w in (1, 2, 3, 5, 8, 13) if w < nbins], np.uint64)
    full = libculoki.scores.snr_boxcar_3d(folds, widths)
    for thr in (0.5, 3.0, 1e9):
        got = libculoki.scores.snr_boxcar_3d_filter(folds, widths, thr, find_max)
        _check(got, _expected(full, widths.size, thr, find_max),
               f"nbins {nbins} max={find_max} thr={thr}")


@pytest.mark.parametrize("name", ["acc17w", "freq16", "jerk17w", "acc_eta2_bseg"])
@pytest.mark.parametrize("fused,storage", [(True, "f32"), (False, "f32"), (True, "f16")])
def test_ffa_filter_equals_thresholded_scores(name, fused, storage):
    from loki_amd import libculoki

    cfg, _ = _cfgs(name)
    ts_e, ts_v = _series(cfg.nsamps, seed=23)
    widths = np.array([w for w in (1, 2, 4, 8) if w < cfg.nbins], np.uint64)
    ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
    ffa.set_fused(fused)
    ffa.set_fold_storage(storage)
    full = ffa.scores(ts_e, ts_v, widths)
    thr = float(np.quantile(full, 0.97))
    for find_max in (False, True):
        got = ffa.scores_filter(ts_e, ts_v, widths, thr, find_max)
        _check(got, _expected(full, widths.size, thr, find_max),
               f"{name} fused={fused} {storage} max={find_max}")


@pytest.mark.parametrize("name", ["freq16", "acc17w"])
def test_fourier_filter_equals_thresholded_scores(name):
    from loki_amd import libculoki

    cfg, _ = _cfgs(name, use_fourier=True)
    ts_e, ts_v = _series(cfg.nsamps, seed=29)
    widths = np.array([1, 2, 3, 6], np.uint64)
    ffa = libculoki.ffa.FFACOMPLEXCUDA(cfg, device_id=0)
    full = ffa.scores(ts_e, ts_v, widths)
    thr = float(np.quantile(full, 0.95))
    for find_max in (False, True):
        got = ffa.scores_filter(ts_e, ts_v, widths, thr, find_max)
        _check(got, _expected(full, widths.size, thr, find_max), f"fourier {name} {find_max}")


def test_compute_ffa_scores_filter_defaults_to_snr_min():
    from loki_amd import libculoki

    cfg, _ = _cfgs("acc17w")
    ts_e, ts_v = _series(cfg.nsamps, seed=31)
    widths = np.array(cfg.score_widths, np.uint64)
    full, _ = libculoki.ffa.compute_ffa_scores_cuda(ts_e, ts_v, cfg)
    idx, snr, plan = libculoki.ffa.compute_ffa_scores_filter_cuda(ts_e, ts_v, cfg, threshold=2.0)
    _check((idx, snr), _expected(full, widths.size, 2.0, False), "compute (2.0)")
    idx, snr, _ = libculoki.ffa.compute_ffa_scores_filter_cuda(ts_e, ts_v, cfg)
    _check((idx, snr), _expected(full, widths.size, cfg.snr_min, False), "compute (snr_min)")


def test_device_filter_overflow_counts_and_writes_capacity():
    import torch

    import loki_amd
    from loki_amd import libculoki

    lib = ctypes.CDLL(str(loki_amd.LIBRARY_PATH))
    fn = lib.loki_snr_boxcar_3d_filter_device
    fn.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p,
                   ctypes.c_size_t, ctypes.c_float, ctypes.c_int, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    lib.loki_last_error.restype = ctypes.c_char_p
    rng = np.random.default_rng(5)
    nprof, nbins = 4096, 32
    folds = np.concatenate([rng.standard_normal((nprof, 1, nbins)) * 3,
                            np.ones((nprof, 1, nbins))], axis=1).astype(np.float32)
    widths = np.array([1, 2, 4], np.uint64)
    full = libculoki.scores.snr_boxcar_3d(folds, widths)
    want_i, want_s = _expected(full, 3, 0.0, False)
    assert want_i.size > 1000
    dev = torch.device("cuda:0")
    tf = torch.from_numpy(folds).to(dev)
    cap = 777
    ti = torch.zeros(cap, dtype=torch.int32, device=dev)
    ts = torch.zeros(cap, dtype=torch.float32, device=dev)
    tc = torch.full((1,), 12345, dtype=torch.int32, device=dev)  # the call zeroes it
    for _ in range(2):  # a second call on the same counter starts from zero again
        rc = fn(tf.data_ptr(), nprof, nbins, widths.ctypes.data, 3, 0.0, 0, ti.data_ptr(),
                ts.data_ptr(), cap, tc.data_ptr(), None)
        assert rc == 0, lib.loki_last_error()
        torch.cuda.synchronize()
        assert int(tc.item()) == want_i.size
    gi = ti.cpu().numpy().view(np.uint32)
    gs = ts.cpu().numpy()
    pos = np.searchsorted(want_i, gi)
    assert np.all(want_i[pos] == gi) and np.unique(gi).size == cap
    np.testing.assert_array_equal(want_s[pos], gs)
    # capacity 0 only counts
    rc = fn(tf.data_ptr(), nprof, nbins, widths.ctypes.data, 3, 0.0, 1, None, None, 0,
            tc.data_ptr(), None)
    assert rc == 0, lib.loki_last_error()
    torch.cuda.synchronize()
    assert int(tc.item()) == int((full.max(axis=1) > 0.0).sum())


def test_filter_errors():
    from loki_amd import libculoki

    folds = np.ones((4, 2, 16), np.float32)
    with pytest.raises(RuntimeError, match="widths"):
        libculoki.scores.snr_boxcar_3d_filter(folds, np.array([16], np.uint64), 1.0)
    with pytest.raises(RuntimeError):
        libculoki.scores.snr_boxcar_3d_filter(folds[0], np.array([1], np.uint64), 1.0)
