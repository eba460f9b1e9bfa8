"""Batched multi-series FFA (FFABatchCUDA): every series' scores are bit-identical to the oracle
and to the single-series GPU path, with the graph replayed several times (with eager work and
synchronisations in between), after autotuning, and with graphs off."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TSAMP = 64e-6
WIDTHS = np.array([1, 2, 3, 4, 6, 8, 12, 16, 24, 32], np.uint64)


def _batch(nseries=9, seed=900):
    from loki_amd import libloki
    from loki_amd.synth import make_batch_series, make_batch_specs

    specs = make_batch_specs(nseries, seed, log2_len=(15, 18), nbins_choices=(16, 32, 64, 128),
                             period_centre=(0.05, 0.5))
    cfgs, series = [], []
    for sp in specs:
        cfgs.append(libloki.configs.PulsarSearchConfig(
            nsamps=sp["nsamps"], tsamp=TSAMP, nbins=sp["nbins"], eta=1.0,
            param_limits=np.array([[sp["f_lo"], sp["f_hi"]]]), use_fourier=False))
        series.append(make_batch_series(sp, TSAMP))
    return specs, cfgs, series


def _widths(nbins):
    return WIDTHS[WIDTHS < nbins]


def test_batch_matches_oracle_and_single_series(oracle):
    from loki_amd import libculoki
    from oracle.plan import SearchConfig

    specs, cfgs, series = _batch()
    assert len({sp["nbins"] for sp in specs}) >= 3 and len({sp["nsamps"] for sp in specs}) >= 2
    b = libculoki.ffa.FFABatchCUDA(cfgs, widths=WIDTHS)
    for i, (e, v) in enumerate(series):
        b.set_input(i, e, v)
    b.execute()
    info = b.info()
    print("batch", info)
    assert info["nseries"] == len(cfgs)
    for i, (sp, (e, v)) in enumerate(zip(specs, series)):
        got = b.scores(i)
        w = _widths(sp["nbins"])
        np.testing.assert_array_equal(b.widths(i), w)
        ocfg = SearchConfig(sp["nsamps"], TSAMP, sp["nbins"], 1.0, [(sp["f_lo"], sp["f_hi"])],
                            use_fourier=False)
        want, _ = oracle.compute_ffa_scores(e, v, ocfg, widths=w)
        np.testing.assert_array_equal(got, want)
        single, _ = libculoki.ffa.compute_ffa_scores_cuda(e, v, cfgs[i], widths=w)
        np.testing.assert_array_equal(got, single)
    assert sum(int(b.plan(i).ncoords[-1]) for i in range(b.size)) == info["total_trials"]


def test_batch_graph_replays_and_schedules():
    import torch

    from loki_amd import libculoki

    specs, cfgs, series = _batch(nseries=7, seed=1300)
    b = libculoki.ffa.FFABatchCUDA(cfgs, widths=WIDTHS)
    for i, (e, v) in enumerate(series):
        b.set_input(i, e, v)
    b.set_graph(False)
    b.execute()
    eager = [b.scores(i) for i in range(b.size)]
    b.set_graph(True)
    stream = torch.cuda.Stream()
    scratch = torch.zeros(1 << 20, device="cuda:0")
    for rep in range(4):
        # the program's own eager work and synchronisations between replays
        scratch.add_(1.0)
        torch.cuda.synchronize()
        for i in range(b.size):
            ptr, n = b.scores_ptr(i)
            torch.cuda.synchronize()
        b.execute(stream.cuda_stream)
        stream.synchronize()
        for i in range(b.size):
            np.testing.assert_array_equal(b.scores(i), eager[i], err_msg=f"replay {rep}")
    assert float(scratch[0].item()) == 4.0
    chosen = b.autotune(repeats=1, stream=stream.cuda_stream)
    print("autotuned schedule", chosen)
    assert set(chosen) == {"merge_group", "merge_run", "final_group", "final_run"}
    b.execute(stream.cuda_stream)
    stream.synchronize()
    for i in range(b.size):
        np.testing.assert_array_equal(b.scores(i), eager[i])
    b.set_schedule({"merge_group": 1, "merge_run": 1, "final_group": 64, "final_run": 4})
    b.execute()
    for i in range(b.size):
        np.testing.assert_array_equal(b.scores(i), eager[i])


def test_batch_device_inputs_from_torch():
    """Series already resident on the GPU (torch tensors) are copied into the batch's slots on
    the caller's stream; results equal the host-input path."""
    import torch

    from loki_amd import libculoki

    specs, cfgs, series = _batch(nseries=4, seed=2100)
    b = libculoki.ffa.FFABatchCUDA(cfgs, widths=WIDTHS)
    ref = libculoki.ffa.FFABatchCUDA(cfgs, widths=WIDTHS)
    stream = torch.cuda.Stream()
    keep = []
    for i, (e, v) in enumerate(series):
        ref.set_input(i, e, v)
        te = torch.from_numpy(e).to("cuda:0")
        tv = torch.from_numpy(v).to("cuda:0")
        torch.cuda.synchronize()
        keep.append((te, tv))
        b.set_input_device(i, te.data_ptr(), tv.data_ptr(), e.size, stream.cuda_stream)
        pe, pv = b.input_ptrs(i)
        assert pe != 0 and pv != 0
    b.execute(stream.cuda_stream)
    stream.synchronize()
    ref.execute()
    for i in range(b.size):
        np.testing.assert_array_equal(b.scores(i), ref.scores(i))


def test_batch_errors():
    from loki_amd import libculoki, libloki

    acc = libloki.configs.PulsarSearchConfig(
        nsamps=2**16, tsamp=TSAMP, nbins=32, eta=1.0,
        param_limits=np.array([[-1e4, 1e4], [20.0, 25.0]]), use_fourier=False)
    with pytest.raises(RuntimeError, match="frequency-only"):
        libculoki.ffa.FFABatchCUDA([acc])
    odd = libloki.configs.PulsarSearchConfig(
        nsamps=2**16, tsamp=TSAMP, nbins=48, eta=1.0, param_limits=np.array([[20.0, 25.0]]),
        use_fourier=False)
    with pytest.raises(RuntimeError, match="nbins"):
        libculoki.ffa.FFABatchCUDA([odd])
