"""FFAs sharing one device workspace (workspace.hpp:169-206 FFAWorkspaceCUDA and the
FFACUDA(workspace, cfg, device_id) constructor, ffa.hpp:100-153).

Plans of different sizes run one after another on one workspace, which grows to the largest;
every result must equal that of an FFA with its own buffers (bit for bit in the time domain,
and the same bits for the Fourier path too, since the kernels are the same).
"""

from __future__ import annotations

import numpy as np
import pytest

from test_gpu_parity import _cfgs, _series

pytestmark = pytest.mark.gpu


def test_shared_workspace_scores_equal_private_buffers():
    from loki_amd import libculoki

    ws = libculoki.ffa.FFAWorkspaceCUDA(device_id=0)
    names = ["freq16", "acc17w", "jerk17w", "acc17_nb128"]
    shared, alone, series = {}, {}, {}
    for n in names:
        cfg, _ = _cfgs(n)
        series[n] = _series(cfg.nsamps, seed=41)
        widths = np.array([w for w in (1, 2, 4, 8) if w < cfg.nbins], np.uint64)
        shared[n] = (libculoki.ffa.FFACUDA(cfg, 0, workspace=ws), widths)
        alone[n] = libculoki.ffa.FFACUDA(cfg, 0).scores(*series[n], widths)
    for _ in range(2):  # interleaved, twice: each run regrows or reuses the buffers
        for n in names:
            ffa, widths = shared[n]
            np.testing.assert_array_equal(ffa.scores(*series[n], widths), alone[n])
            fold = np.zeros(ffa.plan.buffer_size, np.float32)
            ffa.execute(*series[n], fold)


def test_shared_workspace_fourier():
    from loki_amd import libculoki

    ws = libculoki.ffa.FFAWorkspaceCUDA(device_id=0, buffer_bytes=1 << 20, coord_bytes=1 << 16)
    out = {}
    for n in ["freq16", "acc17w"]:
        cfg, _ = _cfgs(n, use_fourier=True)
        ts = _series(cfg.nsamps, seed=43)
        a = libculoki.ffa.FFACOMPLEXCUDA(cfg, 0, workspace=ws)
        b = libculoki.ffa.FFACOMPLEXCUDA(cfg, 0)
        fa = np.zeros(a.plan.buffer_size, np.complex64)
        fb = np.zeros(b.plan.buffer_size, np.complex64)
        a.execute(*ts, fa)
        b.execute(*ts, fb)
        np.testing.assert_array_equal(fa, fb)
        out[n] = a
    del ws  # the FFAs keep the buffers alive
    cfg, _ = _cfgs("freq16", use_fourier=True)
    ts = _series(cfg.nsamps, seed=43)
    f2 = np.zeros(out["freq16"].plan.buffer_size, np.complex64)
    out["freq16"].execute(*ts, f2)
    assert np.isfinite(f2).all()


def test_shared_workspace_two_streams_ordered():
    """Two FFAs sharing one workspace, enqueued back to back on two different streams with no
    host synchronisation between: the second use waits for the first (an event recorded at
    the end of each use), so both score arrays equal those of private buffers."""
    import torch

    from loki_amd import libculoki

    ws = libculoki.ffa.FFAWorkspaceCUDA(device_id=0)
    dev = torch.device("cuda:0")
    runs = []
    for n, seed in (("acc17w", 51), ("jerk17w", 52)):
        cfg, _ = _cfgs(n)
        e, v = _series(cfg.nsamps, seed=seed)
        widths = np.array([w for w in (1, 2, 4, 8) if w < cfg.nbins], np.uint64)
        want = libculoki.ffa.FFACUDA(cfg, 0).scores(e, v, widths)
        ffa = libculoki.ffa.FFACUDA(cfg, 0, workspace=ws)
        out = torch.empty(want.size, dtype=torch.float32, device=dev)
        runs.append((ffa, torch.from_numpy(e).to(dev), torch.from_numpy(v).to(dev), widths,
                     out, want, torch.cuda.Stream(device=dev)))
    torch.cuda.synchronize()
    for _ in range(3):
        for ffa, de, dv, widths, out, want, st in runs:
            ffa.scores_device(de.data_ptr(), dv.data_ptr(), widths, out.data_ptr(), out.numel(),
                              False, st.cuda_stream)
        torch.cuda.synchronize()
        for ffa, de, dv, widths, out, want, st in runs:
            np.testing.assert_array_equal(out.cpu().numpy(), want)
