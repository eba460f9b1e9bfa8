"""GPU parity of the Fourier-domain path against the FP64 oracle.

The reference computes this path in FP32 with fast-math phasor recurrences (CUDA: __sincosf of
FP64 phases, kernels_cuda.cu:780-853 and 512-659; CPU: xsimd recurrences), whose rounding is not
reproducible; its own tests compare implementations at rtol=0.05, atol=2-5
(tests/python/test_cuda_match.py).  The oracle evaluates the same sums in FP64
(oracle/liboracle.c), so parity here is a stated error bound, far tighter than the reference's:

* brute fold (DFT over one segment of K samples, u = 2^-24):
  |gpu - oracle| <= 4 u (sqrt(min(K, 128)) + sqrt(ceil(K / 128)) + 2) * sum_k |x_k| per
  (segment, frequency, component, harmonic).  The twiddles carry ~1 ulp; the matrix cores sum
  blocks of 128 samples in FP32 and the block sums are added in FP32, so a coherent sum (the
  all-positive ts_v at low f*m) errs by ~u (sqrt(128) + sqrt(K/128)) * L1 in the worst random
  case; the factor 4 is the margin.
* each merge level rotates and adds in FP32 (a few ulp of |tail| + |head|); after L levels the
  final fold satisfies max|gpu - oracle| <= 1e-5 (~170 u) * ||oracle row||_2 per
  (coordinate, component) row; measured <= 2e-6.
* returned to the time domain: the same bound on time rows; S/N: |dS/N| <= 1e-4 + 1e-5 |S/N|
  (measured <= 1e-5).
* real transforms against numpy's FP64 FFT: |d| <= 1e-6 * ||row||_1 (IRFFT scaled by 1/n).
"""

from __future__ import annotations

import numpy as np
import pytest

from test_gpu_parity import CASES, _cfgs, _series

pytestmark = pytest.mark.gpu

TSAMP = 64e-6


def _row_check(got, want, width, tol, label):
    got = np.asarray(got).reshape(-1, width)
    want = np.asarray(want).reshape(-1, width)
    assert got.shape == want.shape
    err = np.abs(got.astype(np.complex128) - want.astype(np.complex128)).max(axis=1)
    scale = np.linalg.norm(want.astype(np.complex128), axis=1)
    ratio = float((err / np.maximum(scale, 1e-30)).max())
    print(f"{label}: max row error / row norm = {ratio:.3e} (bound {tol:.0e})")
    assert np.all(err <= tol * scale + 1e-30), ratio


@pytest.mark.parametrize(("nbins", "segment_len", "t_ref"), [
    (64, 2048, 0.0), (64, 512, 0.0123), (32, 1024, 0.05), (7, 256, 0.0),
    (100, 512, 0.01), (256, 1024, 0.0), (16, 8, 0.0), (64, 4, 0.002),
])
def test_brute_fold_fourier_within_bound(oracle, nbins, segment_len, t_ref):
    from loki_amd import libculoki

    nsamps = 2**15
    ts_e, ts_v = _series(nsamps, seed=nbins + segment_len)
    freqs = np.linspace(3.0, 40.0, 19)
    got = libculoki.fold.compute_brute_fold_fourier_cuda(ts_e, ts_v, freqs, segment_len, nbins,
                                                         TSAMP, t_ref)
    want = oracle.brute_fold_fourier(ts_e, ts_v, freqs, segment_len, nbins, TSAMP, t_ref)
    nbf = nbins // 2 + 1
    nseg = nsamps // segment_len
    assert got.dtype == np.complex64 and got.size == want.size == nseg * freqs.size * 2 * nbf
    # bound per segment and component, proportional to the L1 norm of the segment's samples
    u = 2.0**-24
    k = segment_len
    rel = 4 * u * (np.sqrt(min(k, 128)) + np.sqrt(-(-k // 128)) + 2)
    l1 = np.stack([np.abs(ts_e.reshape(nseg, segment_len)).sum(1),
                   np.abs(ts_v.reshape(nseg, segment_len)).sum(1)], axis=1)  # [nseg, 2]
    l1 = np.broadcast_to(l1[:, None, :, None], (nseg, freqs.size, 2, nbf))
    err = np.abs(got.reshape(nseg, freqs.size, 2, nbf).astype(np.complex128)
                 - want.reshape(nseg, freqs.size, 2, nbf).astype(np.complex128))
    print(f"brute fold nbins={nbins} seglen={segment_len}: max err / L1 = "
          f"{float((err / l1).max()):.3e} (bound {rel:.2e})")
    assert np.all(err <= rel * l1)
    # harmonic 0 is the segment sum (exact up to one rounding) with zero imaginary part
    dc = got.reshape(nseg, freqs.size, 2, nbf)[:, :, :, 0]
    assert np.all(dc.imag == 0)


FOURIER_CASES = ["freq16", "acc17w", "acc20", "acc17_nb48", "jerk17w", "jerk20",
                 "acc_eta2_bseg", "freq18_nb100"]


@pytest.mark.parametrize("name", FOURIER_CASES)
def test_ffa_fourier_within_bound(oracle, name):
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name, use_fourier=True)
    ts_e, ts_v = _series(cfg.nsamps, seed=7)
    got, plan = libculoki.ffa.compute_ffa_fourier_cuda(ts_e, ts_v, cfg)
    want, oplan = oracle.compute_ffa(ts_e, ts_v, ocfg, fourier=True)
    assert got.dtype == np.complex64
    assert [list(s) for s in plan.fold_shapes] == oplan.fold_shapes
    assert got.size == plan.fold_size == want.size
    _row_check(got, want, cfg.nbins // 2 + 1, 1e-5, f"fourier {name}")


@pytest.mark.parametrize("name", ["freq16", "acc20", "jerk17w", "freq18_nb100"])
def test_ffa_fourier_return_to_time_within_bound(oracle, name):
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name, use_fourier=True)
    ts_e, ts_v = _series(cfg.nsamps, seed=8)
    got, plan = libculoki.ffa.compute_ffa_fourier_return_to_time_cuda(ts_e, ts_v, cfg)
    fold_c, _ = oracle.compute_ffa(ts_e, ts_v, ocfg, fourier=True)
    want = oracle.irfft(fold_c, cfg.nbins)
    assert got.dtype == np.float32 and got.size == plan.fold_size == want.size
    _row_check(got, want, cfg.nbins, 1e-5, f"return-to-time {name}")


@pytest.mark.parametrize("name", ["freq16", "acc17w", "acc20", "jerk20"])
def test_ffa_scores_fourier_within_bound(oracle, name):
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name, use_fourier=True)
    ts_e, ts_v = _series(cfg.nsamps, seed=9)
    got, plan = libculoki.ffa.compute_ffa_scores_cuda(ts_e, ts_v, cfg)
    want, _ = oracle.compute_ffa_scores(ts_e, ts_v, ocfg)
    assert got.shape == want.shape
    d = np.abs(got.astype(np.float64) - want)
    print(f"scores {name}: max |dS/N| = {float(d.max()):.3e}")
    assert np.all(d <= 1e-4 + 1e-5 * np.abs(want))


@pytest.mark.parametrize("name", ["freq16", "acc17w", "acc_eta2_bseg", "jerk17w"])
@pytest.mark.parametrize(("fused", "exact", "find_max"), [
    (True, True, False), (True, True, True), (True, False, False), (False, True, False),
    (False, False, True)])
def test_fourier_scores_fused_and_unfused(oracle, name, fused, exact, find_max):
    """The fused inverse-DFT + S/N kernel (nbins 16/32/64, matrix cores, no time-domain final
    fold) and the unfused IRFFT + snr_boxcar_3d both stay within the Fourier S/N bound, with
    exact or hardware-rsqrt normalisation, per width and max over widths."""
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name, use_fourier=True)
    ts_e, ts_v = _series(cfg.nsamps, seed=19)
    widths = np.array([w for w in (1, 2, 3, 5, 8, 13) if w < cfg.nbins], dtype=np.uint64)
    ffa = libculoki.ffa.FFACOMPLEXCUDA(cfg, device_id=0)
    ffa.set_fused(fused)
    ffa.set_exact_norm(exact)
    ffa.set_timing(True)
    got = ffa.scores(ts_e, ts_v, widths, find_max)
    kinds = {k for k, _, _ in ffa.timings()}
    assert ("transform_score" in kinds) == (fused and cfg.nbins in (16, 32, 64)), kinds
    want, _ = oracle.compute_ffa_scores(ts_e, ts_v, ocfg, widths=widths)
    want = want.reshape(-1, widths.size)
    if find_max:
        want = want.max(axis=1)
    got = got.reshape(want.shape)
    d = np.abs(got.astype(np.float64) - want)
    print(f"{name} fused={fused} exact={exact} max={find_max}: max |dS/N| = {d.max():.3e}")
    assert np.all(d <= 1e-4 + 1e-5 * np.abs(want))


def test_ffacomplexcuda_class_overloads(oracle):
    """FFACOMPLEXCUDA.execute: complex64 fold of buffer_size, or float fold of 2*buffer_size
    returned to the time domain (loki_templates_cuda.cuh:21-51)."""
    from loki_amd import libculoki

    cfg, ocfg = _cfgs("acc17w", use_fourier=True)
    ts_e, ts_v = _series(cfg.nsamps, seed=10)
    ffa = libculoki.ffa.FFACOMPLEXCUDA(cfg, device_id=0)
    fold_c = np.zeros(ffa.plan.buffer_size, dtype=np.complex64)
    ffa.execute(ts_e, ts_v, fold_c)
    fold_t = np.zeros(2 * ffa.plan.buffer_size, dtype=np.float32)
    ffa.execute(ts_e, ts_v, fold_t)
    want_c, _ = oracle.compute_ffa(ts_e, ts_v, ocfg, fourier=True)
    _row_check(fold_c[: ffa.plan.fold_size], want_c, cfg.nbins // 2 + 1, 1e-5, "class complex")
    want_t = oracle.irfft(want_c, cfg.nbins)
    _row_check(fold_t[: want_t.size], want_t, cfg.nbins, 1e-5, "class time")
    with pytest.raises(RuntimeError, match="buffer_size"):
        ffa.execute(ts_e, ts_v, np.zeros(ffa.plan.buffer_size, dtype=np.float32))


def _capi():
    import ctypes

    import loki_amd

    lib = ctypes.CDLL(str(loki_amd.LIBRARY_PATH))
    for fn in (lib.loki_rfft_batch_device, lib.loki_irfft_batch_device):
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                       ctypes.c_void_p]
        fn.restype = ctypes.c_int
    lib.loki_last_error.restype = ctypes.c_char_p

    def check(rc):
        assert rc == 0, lib.loki_last_error().decode()

    return lib, check


@pytest.mark.parametrize("nbins", [64, 32, 7, 100, 1024])
def test_irfft_rfft_match_numpy(nbins):
    import torch

    lib, check = _capi()
    rng = np.random.default_rng(nbins)
    nrows = 777
    x = rng.standard_normal((nrows, nbins)).astype(np.float32)
    X = np.fft.rfft(x.astype(np.float64), axis=1)
    dev = torch.device("cuda:0")
    tx = torch.from_numpy(x).to(dev)
    tX = torch.empty((nrows, nbins // 2 + 1), dtype=torch.complex64, device=dev)
    check(lib.loki_rfft_batch_device(tx.data_ptr(), tX.data_ptr(), nrows, nbins, None))
    torch.cuda.synchronize()
    got_X = tX.cpu().numpy()
    l1 = np.abs(x).sum(axis=1, keepdims=True)
    assert np.all(np.abs(got_X - X) <= 1e-6 * l1 + 1e-30)
    Xc = X.astype(np.complex64)
    tXc = torch.from_numpy(Xc).to(dev)
    ty = torch.empty((nrows, nbins), dtype=torch.float32, device=dev)
    check(lib.loki_irfft_batch_device(tXc.data_ptr(), ty.data_ptr(), nrows, nbins, None))
    torch.cuda.synchronize()
    want_y = np.fft.irfft(Xc.astype(np.complex128), n=nbins, axis=1)
    l1c = np.abs(Xc).sum(axis=1, keepdims=True) * 2.0 / nbins
    assert np.all(np.abs(ty.cpu().numpy() - want_y) <= 1e-6 * l1c + 1e-30)


def test_fourier_errors():
    from loki_amd import libculoki

    cfg, _ = _cfgs("acc17w", use_fourier=True)
    ts_e, ts_v = _series(cfg.nsamps)
    with pytest.raises(RuntimeError, match="segment length"):
        libculoki.fold.compute_brute_fold_fourier_cuda(ts_e, ts_v, np.array([5.0]), 3000, 64,
                                                       TSAMP)
    with pytest.raises(RuntimeError, match="nbins"):
        libculoki.fold.compute_brute_fold_fourier_cuda(ts_e, ts_v, np.array([5.0]), 1024, 512,
                                                       TSAMP)
    with pytest.raises(RuntimeError, match="nsamps"):
        libculoki.ffa.compute_ffa_fourier_cuda(ts_e[:100], ts_v[:100], cfg)


@pytest.mark.parametrize("name", ["freq16", "acc17w", "acc20", "jerk17w", "freq18_nb100"])
def test_split_bf16_dft_within_bound(oracle, name):
    """The split-bf16 matrix-core brute fold (set_dft_precision("bf16x3"): each operand as a
    high and a low bf16 part, three products per term, FP32 accumulation; ~2^-17 relative per
    product) keeps the final Fourier fold within the FP32 path's bound, 1e-5 of the row norm,
    and the S/N within 1e-4 + 1e-5 |S/N|."""
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name, use_fourier=True)
    ts_e, ts_v = _series(cfg.nsamps, seed=7)
    ffa = libculoki.ffa.FFACOMPLEXCUDA(cfg, device_id=0)
    ffa.set_dft_precision("bf16x3")
    fold_c = np.zeros(ffa.plan.buffer_size, dtype=np.complex64)
    ffa.execute(ts_e, ts_v, fold_c)
    want, _ = oracle.compute_ffa(ts_e, ts_v, ocfg, fourier=True)
    _row_check(fold_c[: ffa.plan.fold_size], want, cfg.nbins // 2 + 1, 1e-5, f"bf16x3 {name}")
    widths = np.array([w for w in (1, 2, 3, 5, 8, 13) if w < cfg.nbins], dtype=np.uint64)
    got = ffa.scores(ts_e, ts_v, widths)
    want_s, _ = oracle.compute_ffa_scores(ts_e, ts_v, ocfg, widths=widths)
    d = np.abs(got.astype(np.float64) - want_s)
    print(f"bf16x3 scores {name}: max |dS/N| = {d.max():.3e}")
    assert np.all(d <= 1e-4 + 1e-5 * np.abs(want_s))
    with pytest.raises(ValueError):
        ffa.set_dft_precision("f16")
