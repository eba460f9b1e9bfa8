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
    print("batch built, executing", flush=True)
    b.execute()
    assert b.errors() == 0
    info = b.info()
    print("batch", info, flush=True)
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
    assert set(chosen) == {"merge_group", "merge_run", "final_group", "final_run", "persistent",
                           "score_lanes_8", "score_lanes_16", "score_lanes_32", "score_lanes_64",
                           "score_lanes_128"}
    b.execute(stream.cuda_stream)
    stream.synchronize()
    for i in range(b.size):
        np.testing.assert_array_equal(b.scores(i), eager[i])
    b.set_schedule({"merge_group": 1, "merge_run": 1, "final_group": 64, "final_run": 4})
    b.execute()
    for i in range(b.size):
        np.testing.assert_array_equal(b.scores(i), eager[i])


@pytest.mark.skip(reason="the persistent launch (k_tree.hip) stalled on MI355X with no fault: "
                         "no output for 150 s in its first launch, twice (acquire polling, then "
                         "relaxed polling); the mode is disabled (set_schedule rejects it)")
def test_batch_persistent_launch_equals_grouped_launches():
    """One persistent launch (every stage of every series, tiles waiting on per-item counters)
    gives the grouped launches' bits for every lanes-per-trial choice of the last levels, over
    eager runs and graph replays, and never flags a given-up wait."""
    import torch

    from loki_amd import libculoki

    specs, cfgs, series = _batch(nseries=10, seed=3100)
    assert {sp["nbins"] for sp in specs} >= {32, 64, 128}
    b = libculoki.ffa.FFABatchCUDA(cfgs, widths=WIDTHS)
    for i, (e, v) in enumerate(series):
        b.set_input(i, e, v)
    b.set_graph(False)
    b.set_schedule({"persistent": 0})
    b.execute()
    grouped = [b.scores(i) for i in range(b.size)]
    assert b.info()["launches"] > 1
    schedules = [{"persistent": 1},
                 {"persistent": 1, "score_lanes_16": 4, "score_lanes_32": 2,
                  "score_lanes_64": 4, "score_lanes_128": 4},
                 {"persistent": 1, "score_lanes_8": 2, "score_lanes_16": 2, "score_lanes_32": 4,
                  "score_lanes_64": 2, "merge_group": 1, "merge_run": 64, "final_group": 64,
                  "final_run": 4}]
    stream = torch.cuda.Stream()
    for sc in schedules:
        print("schedule", sc, flush=True)
        b.set_schedule(sc)
        assert b.info()["launches"] == 1
        b.set_graph(False)
        b.execute()
        assert b.errors() == 0
        for i in range(b.size):
            np.testing.assert_array_equal(b.scores(i), grouped[i], err_msg=f"{sc} series {i}")
        b.set_graph(True)
        for rep in range(3):
            torch.cuda.synchronize()
            b.execute(stream.cuda_stream)
            stream.synchronize()
            assert b.errors() == 0
            for i in range(b.size):
                np.testing.assert_array_equal(b.scores(i), grouped[i],
                                              err_msg=f"{sc} replay {rep} series {i}")


def test_batch_width_schedules_equal_one_lane_per_trial():
    """The boxcar-width schedules of the last levels (1, 2 or 4 lanes per trial in 256-lane
    workgroups, lane q scoring widths q, q + lanes, ...) give the bits of the 64-lane kernel
    (one lane per trial, all widths), eagerly and over graph replays."""
    import torch

    from loki_amd import libculoki

    specs, cfgs, series = _batch(nseries=10, seed=3100)
    assert {sp["nbins"] for sp in specs} >= {16, 32, 64, 128}
    b = libculoki.ffa.FFABatchCUDA(cfgs, widths=WIDTHS)
    for i, (e, v) in enumerate(series):
        b.set_input(i, e, v)
    b.set_graph(False)
    b.execute()
    ref = [b.scores(i) for i in range(b.size)]
    schedules = [{"score_lanes_16": 1, "score_lanes_32": 1, "score_lanes_64": 2,
                  "score_lanes_128": 4},
                 {"score_lanes_16": 2, "score_lanes_32": 2, "score_lanes_64": 4,
                  "score_lanes_128": 4, "final_group": 4, "final_run": 64},
                 {"score_lanes_16": 4, "score_lanes_32": 4, "score_lanes_64": 2,
                  "score_lanes_128": 4, "final_group": 64, "final_run": 4}]
    stream = torch.cuda.Stream()
    for sc in schedules:
        print("schedule", sc, flush=True)
        b.set_schedule(sc)
        b.set_graph(False)
        b.execute()
        assert b.errors() == 0
        for i in range(b.size):
            np.testing.assert_array_equal(b.scores(i), ref[i], err_msg=f"{sc} series {i}")
        b.set_graph(True)
        for rep in range(2):
            torch.cuda.synchronize()
            b.execute(stream.cuda_stream)
            stream.synchronize()
            for i in range(b.size):
                np.testing.assert_array_equal(b.scores(i), ref[i],
                                              err_msg=f"{sc} replay {rep} series {i}")


def test_batch_width_schedules_small_nbins_and_wide_windows(oracle):
    """Width schedules at the edges: nbins 8 and 16 (two lanes per trial at most for 8) and
    nbins 64 with windows up to 63 bins, whose prefix rows need four lanes per trial to fit the
    64 KiB workgroup budget (two lanes are rejected); every score equals the oracle's."""
    from loki_amd import libculoki, libloki
    from loki_amd.synth import make_batch_series, make_batch_specs
    from oracle.plan import SearchConfig

    specs = make_batch_specs(6, 4100, log2_len=(15, 17), nbins_choices=(8, 16, 64),
                             period_centre=(0.05, 0.3))
    assert {sp["nbins"] for sp in specs} == {8, 16, 64}
    widths = np.arange(1, 64, dtype=np.uint64)
    cfgs, series = [], []
    for sp in specs:
        cfgs.append(libloki.configs.PulsarSearchConfig(
            nsamps=sp["nsamps"], tsamp=TSAMP, nbins=sp["nbins"], eta=1.0,
            param_limits=np.array([[sp["f_lo"], sp["f_hi"]]]), use_fourier=False))
        series.append(make_batch_series(sp, TSAMP))
    b = libculoki.ffa.FFABatchCUDA(cfgs, widths=widths)
    for i, (e, v) in enumerate(series):
        b.set_input(i, e, v)
    with pytest.raises(RuntimeError, match="64 KiB"):
        b.set_schedule({"score_lanes_64": 2})
    b.set_graph(False)
    for sc in ({"score_lanes_8": 1, "score_lanes_16": 4, "score_lanes_64": 4},
               {"score_lanes_8": 2, "score_lanes_16": 2, "score_lanes_64": 4}):
        print("schedule", sc, flush=True)
        b.set_schedule(sc)
        b.execute()
        for i, (sp, (e, v)) in enumerate(zip(specs, series)):
            w = widths[widths < sp["nbins"]]
            ocfg = SearchConfig(sp["nsamps"], TSAMP, sp["nbins"], 1.0, [(sp["f_lo"], sp["f_hi"])],
                                use_fourier=False)
            want, _ = oracle.compute_ffa_scores(e, v, ocfg, widths=w)
            np.testing.assert_array_equal(b.scores(i), want, err_msg=f"{sc} series {i}")


def test_batch_schedule_errors():
    from loki_amd import libculoki

    _, cfgs, _ = _batch(nseries=3, seed=3300)
    b = libculoki.ffa.FFABatchCUDA(cfgs, widths=WIDTHS)
    for bad in ({"score_lanes_8": 4}, {"score_lanes_128": 2}, {"score_lanes_32": 3},
                {"score_lanes_12": 1}, {"merge_group": 0}, {"persistent": 1}):
        with pytest.raises(RuntimeError):
            b.set_schedule(bad)


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
