"""Host-side C-ABI of the region planner, the frequency sweep and the seed entry points (CPU).

No device work: the planner and the sweep's constructor are host code, and the seed entry
points reject bad arguments before any launch.  Also the bench contract's refusal of more
than one GPU.
"""

from __future__ import annotations

import ctypes
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


class Region(ctypes.Structure):
    _fields_ = [("f_start", ctypes.c_double), ("f_end", ctypes.c_double),
                ("nbins", ctypes.c_size_t), ("eta", ctypes.c_double)]


@pytest.fixture(scope="module")
def lib():
    import loki_amd

    L = ctypes.CDLL(str(loki_amd.LIBRARY_PATH))
    L.loki_last_error.restype = ctypes.c_char_p
    L.loki_generate_ffa_regions.argtypes = [ctypes.c_double] * 3 + [ctypes.c_size_t] + \
        [ctypes.c_double] * 2 + [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                 ctypes.POINTER(ctypes.c_size_t)]
    return L


def test_generate_regions_query_then_fill(lib):
    n = ctypes.c_size_t(0)
    assert lib.loki_generate_ffa_regions(1 / 500.0, 1.0, 64e-6, 64, 1.0, 2.0, 0, None, 0,
                                         ctypes.byref(n)) == 0
    assert n.value == 9
    arr = (Region * n.value)()
    assert lib.loki_generate_ffa_regions(1 / 500.0, 1.0, 64e-6, 64, 1.0, 2.0, 0, arr, n.value,
                                         ctypes.byref(n)) == 0
    assert [r.nbins for r in arr] == [31 * 2**k for k in range(9)]
    small = (Region * 2)()
    assert lib.loki_generate_ffa_regions(1 / 500.0, 1.0, 64e-6, 64, 1.0, 2.0, 0, small, 2,
                                         ctypes.byref(n)) != 0
    assert b"too small" in lib.loki_last_error()
    assert lib.loki_generate_ffa_regions(1.0, 0.5, 64e-6, 64, 1.0, 2.0, 0, None, 0,
                                         ctypes.byref(n)) != 0
    assert b"p_max" in lib.loki_last_error()


def test_planner_and_sweep_handles(lib):
    from loki_amd import libculoki, libloki

    cfg = libloki.configs.PulsarSearchConfig(
        nsamps=2**18, tsamp=1e-4, nbins=32, eta=1.0, param_limits=np.array([[-5.0, 5.0],
                                                                           [20.0, 90.0]]),
        use_fourier=False, max_passing_candidates=1 << 12)
    sweep = libculoki.ffa.FFAFreqSweepCUDA(cfg)
    pl = sweep.planner
    direct = libloki.plans.FFARegionPlannerTime(cfg)
    assert pl.nregions == direct.nregions == 3
    assert [c.nbins for c in pl.cfgs] == [c.nbins for c in direct.cfgs]
    snr, rows = sweep.results()  # nothing run yet
    assert snr.size == 0 and rows.shape == (0, 3)
    assert list(sweep.per_region) == [0, 0, 0]
    with pytest.raises(RuntimeError, match="fold kind|fold_type|kind"):
        libloki.plans.FFARegionPlanner(cfg, fold_type=7)


def test_seed_entry_points_reject_bad_sizes(lib):
    vp, sz, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32
    f = lib.loki_shift_add_linear_complex_batch_device
    f.argtypes = [vp] * 7 + [sz, sz, sz, u32, u32, sz, vp]
    dummy = ctypes.c_void_p(16)  # never dereferenced: the sizes are rejected first
    assert f(dummy, dummy, None, dummy, dummy, dummy, dummy, 8, 16, 4, 0, 4, 4, None) != 0
    assert b"nbins_f" in lib.loki_last_error()
    g = lib.loki_shift_add_ascend_linear_batch_device
    g.argtypes = [vp] * 5 + [sz] * 5 + [vp]
    assert g(dummy, dummy, dummy, dummy, dummy, 0, 4, 4, 1, 1, None) != 0
    assert b"nbins" in lib.loki_last_error()
    assert g(None, None, None, None, None, 16, 4, 4, 1, 1, None) != 0
    assert b"null" in lib.loki_last_error()


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
