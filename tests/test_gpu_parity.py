"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Bar: bit-exact for shift/index tables, time-domain folds and boxcar S/N (the oracle fixes the
reference CPU's operation order, see oracle/liboracle.c).  Mirrors the structure of the
reference's tests/python/test_cuda_match.py (brute fold, freq / accel / jerk FFA, scores), with
seeded inputs and exact comparisons instead of decimal=3.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TSAMP = 64e-6

CASES = {
    # name: (nsamps, nbins, param_limits, extra kwargs)
    "freq16": (2**16, 64, [(2.0, 8.0)], {}),
    "freq18": (2**18, 64, [(1.0, 3.0)], {}),
    "freq18_nb100": (2**18, 100, [(1.0, 3.0)], {}),
    "acc17w": (2**17, 32, [(-60000.0, 60000.0), (100.0, 110.0)], {}),
    "acc20": (2**20, 64, [(-1000.0, 1000.0), (100.0, 102.0)], {}),
    "acc17_nb48": (2**17, 48, [(-60000.0, 60000.0), (100.0, 110.0)], {}),
    "jerk17w": (2**17, 32, [(-2e5, 2e5), (-6e4, 6e4), (50.0, 55.0)], {}),
    "jerk20": (2**20, 32, [(-50.0, 50.0), (-500.0, 500.0), (60.0, 61.0)], {}),
    "acc_eta2_bseg": (2**18, 16, [(-20000.0, 20000.0), (30.0, 40.0)],
                      {"eta": 2.0, "bseg_brute": 2048}),
    "acc17_nb128": (2**17, 128, [(-60000.0, 60000.0), (60.0, 64.0)], {}),
    # the one-lane-per-trial fused kernels (nbins 8, 16, 32): one grid row, 36 rows in
    # (jerk, accel), and trees of 2 and 3 levels (the parent level is the brute fold, or is
    # merged on chip from it)
    "freq17_nb32": (2**17, 32, [(20.0, 30.0)], {}),
    "acc16_nb8": (2**16, 8, [(-300000.0, 300000.0), (200.0, 230.0)], {}),
    "jerk18_nb16": (2**18, 16, [(-1e5, 1e5), (-3e4, 3e4), (40.0, 44.0)], {}),
    "acc_2lvl": (2**16, 32, [(-60000.0, 60000.0), (100.0, 110.0)], {"bseg_brute": 2**15}),
    "acc_3lvl": (2**16, 32, [(-60000.0, 60000.0), (100.0, 110.0)], {"bseg_brute": 2**14}),
}


def _cfgs(name, use_fourier=False):
    from loki_amd import libloki
    from oracle.plan import SearchConfig

    nsamps, nbins, lims, extra = CASES[name]
    eta = extra.get("eta", 1.0)
    kw = {k: v for k, v in extra.items() if k != "eta"}
    cfg = libloki.configs.PulsarSearchConfig(
        nsamps=nsamps, tsamp=TSAMP, nbins=nbins, eta=eta, param_limits=np.array(lims),
        use_fourier=use_fourier, **kw)
    ocfg = SearchConfig(nsamps, TSAMP, nbins, eta, lims, use_fourier=use_fourier, **kw)
    return cfg, ocfg


def _series(nsamps, seed=1234):
    rng = np.random.default_rng(seed)
    ts_e = rng.standard_normal(nsamps, dtype=np.float32)
    ts_v = (1.0 + 0.25 * rng.random(nsamps)).astype(np.float32)
    return ts_e, ts_v


@pytest.mark.parametrize("name", list(CASES))
def test_plan_geometry_matches_oracle(name):
    from loki_amd import libloki
    from oracle.plan import Plan

    cfg, ocfg = _cfgs(name)
    plan = libloki.plans.FFAPlanTime(cfg)
    oplan = Plan(ocfg)
    assert plan.n_levels == oplan.n_levels
    assert [list(s) for s in plan.fold_shapes] == oplan.fold_shapes
    assert list(plan.ncoords) == oplan.ncoords
    assert list(plan.ncoords_offsets) == oplan.ncoords_offsets
    assert plan.buffer_size == oplan.buffer_size
    for lvl in range(plan.n_levels):
        np.testing.assert_array_equal(plan.dparams[lvl], np.array(oplan.dparams[lvl]))
        for a, b in zip(plan.compute_param_grid(lvl), oplan.compute_param_grid(lvl)):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name", list(CASES))
def test_resolve_coordinates_bit_exact(name):
    """FP64 HIP resolver vs FP64 oracle (taylor_ffa.cpp:19-311): identical bits."""
    from loki_amd import libloki
    from oracle.plan import Plan

    cfg, ocfg = _cfgs(name)
    plan = libloki.plans.FFAPlanTime(cfg)
    oplan = Plan(ocfg)
    if oplan.n_params == 1:
        got = plan.resolve_coordinates_freq()
    else:
        got = plan.resolve_coordinates()
    want = oplan.resolve_coordinates()
    assert len(got) == len(want)
    for lvl, (g, w) in enumerate(zip(got, want)):
        assert g.dtype.names == w.dtype.names
        for field in w.dtype.names:
            np.testing.assert_array_equal(g[field], w[field], err_msg=f"level {lvl} {field}")


@pytest.mark.parametrize(
    ("nbins", "segment_len", "t_ref"),
    [(64, 2**13, 0.0), (32, 1024, 0.0328), (100, 4096, 0.0), (128, 2**14, 0.5), (7, 2048, 0.1)],
)
def test_brute_fold_time_bit_exact(oracle, nbins, segment_len, t_ref):
    _check_brute_fold(oracle, nbins, segment_len, t_ref, np.linspace(140.0, 145.0, 50))


# Segments longer than 4096 samples take the run path (contiguous runs of one phase bin,
# boundaries found from the exact FP64 bin function): long periods, periods shorter than nbins
# (empty runs), less than one period per segment, and negative phases (t_ref > 0).
@pytest.mark.parametrize(("nbins", "segment_len", "t_ref", "f_lo", "f_hi", "nf"), [
    (64, 2**16, 0.0, 0.5, 2.0, 23), (128, 2**15, 1.1, 0.7, 1.3, 9), (32, 2**13, 0.0, 3000.0,
    5000.0, 7), (64, 2**16, 0.0, 0.01, 0.05, 5), (16, 2**17, 0.37, 9.9, 10.1, 5),
    (100, 8192, 0.0, 1.0, 40.0, 11),
])
def test_brute_fold_time_run_path_bit_exact(oracle, nbins, segment_len, t_ref, f_lo, f_hi, nf):
    _check_brute_fold(oracle, nbins, segment_len, t_ref, np.linspace(f_lo, f_hi, nf))


# Segments of at most 2048 samples are folded several per block (8, 4 or 2 as the segment
# length allows): segment counts that leave the last block of segments partly empty.
@pytest.mark.parametrize(("nbins", "segment_len", "nsamps", "t_ref"), [
    (32, 512, 13 * 512, 0.0), (64, 1024, 9 * 1024, 0.21), (37, 2000, 5 * 2000, 0.0),
    (16, 300, 11 * 300, 0.05), (8, 64, 3 * 64, 0.0), (32, 1024, 1024, 0.0),
])
def test_brute_fold_time_ragged_segment_groups(oracle, nbins, segment_len, nsamps, t_ref):
    _check_brute_fold(oracle, nbins, segment_len, t_ref, np.linspace(140.0, 145.0, 37),
                      nsamps=nsamps)


def _check_brute_fold(oracle, nbins, segment_len, t_ref, freqs, nsamps=2**17):
    """Brute fold vs the ascending-order bucket sums (fold.cpp:113-142, kernels.cpp:550-580)."""
    from loki_amd import libculoki

    ts_e, ts_v = _series(nsamps, seed=7)
    got = libculoki.fold.compute_brute_fold_time_cuda(
        ts_e, ts_v, freqs, segment_len=segment_len, nbins=nbins, tsamp=TSAMP, t_ref=t_ref,
        device_id=0)
    want = oracle.brute_fold_time(ts_e, ts_v, freqs, segment_len, nbins, TSAMP, t_ref)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("name", list(CASES))
def test_ffa_time_bit_exact(oracle, name):
    """compute_ffa_time_cuda vs the oracle level loop: identical fold cube."""
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name)
    ts_e, ts_v = _series(cfg.nsamps, seed=11)
    got, plan = libculoki.ffa.compute_ffa_time_cuda(ts_e, ts_v, cfg, device_id=0, quiet=True)
    want, oplan = oracle.compute_ffa(ts_e, ts_v, ocfg)
    assert got.size == plan.fold_size == want.size
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("name", ["freq16", "freq18", "acc17w", "acc20", "jerk17w", "jerk20"])
def test_ffa_scores_bit_exact(oracle, name):
    """compute_ffa_scores_cuda vs oracle FFA + snr_boxcar_3d (ffa.cpp:444-466)."""
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name)
    ts_e, ts_v = _series(cfg.nsamps, seed=3)
    got, plan = libculoki.ffa.compute_ffa_scores_cuda(ts_e, ts_v, cfg, device_id=0, quiet=True)
    want, _ = oracle.compute_ffa_scores(ts_e, ts_v, ocfg)
    np.testing.assert_array_equal(got, want)
    # explicit widths (the benchmark's list minus widths >= nbins)
    widths = np.array([w for w in (1, 2, 3, 4, 6, 8, 12, 16, 24, 32) if w < cfg.nbins],
                      dtype=np.uint64)
    got2, _ = libculoki.ffa.compute_ffa_scores_cuda(ts_e, ts_v, cfg, widths=widths)
    want2, _ = oracle.compute_ffa_scores(ts_e, ts_v, ocfg, widths=widths)
    np.testing.assert_array_equal(got2, want2)


@pytest.mark.parametrize("name", ["freq16", "acc17w", "acc20", "jerk17w", "jerk20",
                                  "freq18_nb100", "acc_eta2_bseg", "acc17_nb128", "acc17_nb48",
                                  "freq17_nb32", "acc16_nb8", "jerk18_nb16", "acc_2lvl",
                                  "acc_3lvl"])
@pytest.mark.parametrize("mode", ["unfused", "depth1", "depth2"])
@pytest.mark.parametrize("find_max", [False, True])
def test_scores_fused_and_unfused_paths(oracle, name, mode, find_max):
    """The fused last-level kernels (FP64 shift tables + merge + S/N, no final fold in HBM;
    depth 2 also merges the parent level on chip) and the unfused merge + snr_boxcar_3d give
    the oracle's scores bit for bit."""
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name)
    ts_e, ts_v = _series(cfg.nsamps, seed=17)
    widths = np.array([w for w in (1, 2, 3, 5, 8, 13) if w < cfg.nbins], dtype=np.uint64)
    ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
    ffa.set_fused(mode != "unfused")
    if mode != "unfused":
        ffa.set_fuse_depth(2 if mode == "depth2" else 1)
    got = ffa.scores(ts_e, ts_v, widths, find_max=find_max)
    want, _ = oracle.compute_ffa_scores(ts_e, ts_v, ocfg, widths=widths)
    want = want.reshape(-1, widths.size)
    np.testing.assert_array_equal(got, want.max(axis=1) if find_max else want.reshape(-1))


@pytest.mark.parametrize("name", ["freq16", "acc20", "jerk17w"])
def test_scores_fast_norm_within_tolerance(oracle, name):
    """Hardware-rsqrt normalisation: S/N within 1e-4 of the oracle (absolute and relative)."""
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name)
    ts_e, ts_v = _series(cfg.nsamps, seed=23)
    widths = np.array([w for w in (1, 2, 4, 8, 16) if w < cfg.nbins], dtype=np.uint64)
    ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
    ffa.set_exact_norm(False)
    got = ffa.scores(ts_e, ts_v, widths)
    want, _ = oracle.compute_ffa_scores(ts_e, ts_v, ocfg, widths=widths)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("nbins", [8, 16, 32, 64, 128, 100, 256])
def test_snr_boxcar_3d_bit_exact(oracle, nbins):
    from loki_amd import libculoki

    rng = np.random.default_rng(nbins)
    nprof = 1000
    e = rng.standard_normal((nprof, 1, nbins)).astype(np.float32) * 30
    v = (rng.random((nprof, 1, nbins)) * 100 + 1).astype(np.float32)
    folds = np.ascontiguousarray(np.concatenate([e, v], axis=1))
    widths = np.array([w for w in (1, 2, 3, 4, 6, 9, 13, 19, 31, 63, 99) if w < nbins],
                      dtype=np.uint64)
    got = libculoki.scores.snr_boxcar_3d(folds, widths, device_id=0)
    want = oracle.snr_boxcar_3d(folds, widths)
    np.testing.assert_array_equal(got, want)
    got_max = libculoki.scores.snr_boxcar_3d_max(folds, widths, device_id=0)
    np.testing.assert_array_equal(got_max, want.max(axis=1))


def test_snr_boxcar_2d_max(oracle):
    from loki_amd import libculoki

    rng = np.random.default_rng(5)
    arr = rng.standard_normal((300, 64)).astype(np.float32)
    widths = np.array([1, 2, 4, 8], dtype=np.uint64)
    got = libculoki.scores.snr_boxcar_2d_max(arr, widths, stdnoise=2.0)
    folds = np.stack([arr, np.ones_like(arr)], axis=1)
    want = oracle.snr_boxcar_3d(folds, widths).max(axis=1) * np.float32(0.5)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


def test_ffa_execute_device_torch_stream(oracle):
    """Device-pointer entry point on a torch-owned buffer and stream (async path)."""
    import torch

    from loki_amd import libculoki

    cfg, ocfg = _cfgs("acc17w")
    ts_e, ts_v = _series(cfg.nsamps, seed=21)
    ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
    dev = torch.device("cuda:0")
    te = torch.from_numpy(ts_e).to(dev)
    tv = torch.from_numpy(ts_v).to(dev)
    fold = torch.empty(ffa.plan.buffer_size, dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(stream):
        ffa.execute_device(te.data_ptr(), tv.data_ptr(), fold.data_ptr(), fold.numel(),
                           stream.cuda_stream)
    stream.synchronize()
    got = fold[: ffa.plan.fold_size].cpu().numpy()
    want, _ = oracle.compute_ffa(ts_e, ts_v, ocfg)
    np.testing.assert_array_equal(got, want)
    # repeat run on the same object: identical (deterministic, no atomics)
    with torch.cuda.stream(stream):
        ffa.execute_device(te.data_ptr(), tv.data_ptr(), fold.data_ptr(), fold.numel(),
                           stream.cuda_stream)
    stream.synchronize()
    np.testing.assert_array_equal(fold[: ffa.plan.fold_size].cpu().numpy(), want)


def test_ffacuda_class_execute_host(oracle):
    from loki_amd import libculoki

    cfg, ocfg = _cfgs("jerk17w")
    ts_e, ts_v = _series(cfg.nsamps, seed=9)
    ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
    fold = np.zeros(ffa.plan.buffer_size, dtype=np.float32)
    ffa.execute(ts_e, ts_v, fold)
    want, _ = oracle.compute_ffa(ts_e, ts_v, ocfg)
    np.testing.assert_array_equal(fold[: ffa.plan.fold_size], want)


def test_errors_raise_runtime_error():
    from loki_amd import libculoki

    cfg, _ = _cfgs("freq16")
    ts_e, ts_v = _series(cfg.nsamps)
    ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
    with pytest.raises(RuntimeError, match="buffer_size"):
        ffa.execute(ts_e, ts_v, np.zeros(10, dtype=np.float32))
    with pytest.raises(RuntimeError, match="nsamps"):
        libculoki.ffa.compute_ffa_time_cuda(ts_e[:100], ts_v[:100], cfg)
    with pytest.raises(RuntimeError, match="device_id"):
        libculoki.ffa.FFACUDA(cfg, device_id=99)
    with pytest.raises(RuntimeError, match="widths"):
        libculoki.scores.snr_boxcar_3d(np.ones((4, 2, 8), np.float32), np.array([8], np.uint64))


# ---------------------------------------------------------------------------
# Reduced-precision fold storage: bit-exact against the oracle's emulation
# (same FP32 sums, each stored value rounded once to FP16 / BF16 at the scale
# 2^-(k0 + l), or each component row to block floating point with a shared
# exponent); the S/N error against FP64 is stated in test_storage_snr_error.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("storage", ["f16", "bf16", "bfp16"])
@pytest.mark.parametrize("name", ["freq16", "acc17w", "acc20", "jerk17w", "freq18_nb100",
                                  "acc_eta2_bseg", "acc17_nb128"])
def test_reduced_storage_fold_bit_exact(oracle, name, storage):
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name)
    ts_e, ts_v = _series(cfg.nsamps, seed=31)
    if storage == "bfp16" and cfg.nbins & (cfg.nbins - 1):
        with pytest.raises(RuntimeError, match="power-of-two nbins"):
            libculoki.ffa.compute_ffa_time_cuda(ts_e, ts_v, cfg, fold_storage=storage)
        return
    got, _ = libculoki.ffa.compute_ffa_time_cuda(ts_e, ts_v, cfg, fold_storage=storage)
    want, _ = oracle.compute_ffa(ts_e, ts_v, ocfg, storage=storage)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("storage", ["f16", "bf16", "bfp16"])
@pytest.mark.parametrize("name", ["freq16", "acc17w", "acc20", "jerk17w", "jerk20",
                                  "acc_eta2_bseg", "freq17_nb32", "jerk18_nb16", "acc_3lvl",
                                  "acc16_nb8", "acc17_nb128"])
@pytest.mark.parametrize("mode", ["unfused", "depth1", "depth2"])
def test_reduced_storage_scores_bit_exact(oracle, name, storage, mode):
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name)
    ts_e, ts_v = _series(cfg.nsamps, seed=32)
    widths = np.array([w for w in (1, 2, 3, 4, 6, 8, 12, 16, 24) if w < cfg.nbins], np.uint64)
    ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
    ffa.set_fold_storage(storage)
    ffa.set_fused(mode != "unfused")
    if mode != "unfused":
        ffa.set_fuse_depth(2 if mode == "depth2" else 1)
    got = ffa.scores(ts_e, ts_v, widths)
    want, _ = oracle.compute_ffa_scores(ts_e, ts_v, ocfg, widths=widths, storage=storage)
    np.testing.assert_array_equal(got, want)


# Stated S/N error of each storage precision against the FP64 oracle (FP64 folds and FP64
# scoring), on series with an injected pulsar (S/N 15 at the best width).  FP32 storage is
# ~1e-6 (FP32 rounding only); FP16 keeps 11 significand bits, so a strong candidate carries
# about S/N * 2^-12 per rounding (~5e-3 at S/N 15); BF16 (8 bits) about 8x more.  Block
# floating point keeps 15 bits relative to each row's largest value: within the 1e-3 bar.
STORAGE_SNR_BOUND = {"f32": 2e-5, "f16": 1e-2, "bf16": 8e-2, "bfp16": 1e-3}


@pytest.mark.parametrize("name", ["acc20", "jerk20"])
def test_storage_snr_error_vs_fp64(oracle, name):
    from loki_amd import libculoki
    from loki_amd.synth import make_series

    cfg, ocfg = _cfgs(name)
    nbins = cfg.nbins
    widths = np.array([w for w in (1, 2, 3, 4, 6, 8, 12, 16, 24, 32) if w < nbins], np.uint64)
    lims = CASES[name][2]
    ts_e, ts_v, _ = make_series(cfg.nsamps, TSAMP, 77, f0=lims[-1][0] + 0.37, fwhm=0.05,
                                snr=15.0, nbins=nbins, widths=widths)
    fold64, _ = oracle.compute_ffa(ts_e, ts_v, ocfg, f64=True)
    snr64 = oracle.snr_boxcar_3d_f64(fold64, widths, nbins).reshape(-1)
    ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
    for storage, bound in STORAGE_SNR_BOUND.items():
        ffa.set_fold_storage(storage)
        got = ffa.scores(ts_e, ts_v, widths).astype(np.float64)
        d = np.abs(got - snr64)
        top = np.argsort(snr64)[-100:]
        print(f"{name} {storage}: max |dS/N| {d.max():.2e}, p99 {np.percentile(d, 99):.2e}, "
              f"median {np.median(d):.2e}; top-100 candidates max {d[top].max():.2e} "
              f"(S/N up to {snr64.max():.1f})")
        assert d.max() <= bound


# ---------------------------------------------------------------------------
# Two-level merge passes: the same bits as one pass per level
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16", "bfp16"])
@pytest.mark.parametrize("name", ["freq16", "acc17w", "jerk17w", "acc20", "acc16_nb8",
                                  "jerk18_nb16", "acc17_nb128", "acc_eta2_bseg"])
def test_two_level_merges_equal_single_levels(oracle, name, storage):
    """Folds (every level returned through execute) and fused scores with two-level merge
    passes equal one pass per level, and the oracle."""
    from loki_amd import libculoki

    cfg, ocfg = _cfgs(name)
    ts_e, ts_v = _series(cfg.nsamps, seed=71)
    widths = np.array([w for w in (1, 2, 3, 4, 6, 8, 12, 16, 24) if w < cfg.nbins], np.uint64)
    out = {}
    for two in (True, False):
        ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
        ffa.set_fold_storage(storage)
        ffa.set_two_level_merges(two)
        fold = np.zeros(ffa.plan.buffer_size, np.float32)
        ffa.execute(ts_e, ts_v, fold)
        out[two] = (fold[: ffa.plan.fold_size].copy(), ffa.scores(ts_e, ts_v, widths))
    np.testing.assert_array_equal(out[True][0], out[False][0])
    np.testing.assert_array_equal(out[True][1], out[False][1])
    want, _ = oracle.compute_ffa_scores(ts_e, ts_v, ocfg, widths=widths, storage=storage)
    np.testing.assert_array_equal(out[True][1], want)
