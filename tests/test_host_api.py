"""Host-side API (CPU only): the C-ABI library loads and exports every declared symbol, and the
config / plan classes behind it (C++20, no GPU involved) match the oracle and the reference's
known answers, with the reference's error behaviour (RuntimeError on invalid input)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_functions() -> list[str]:
    text = (ROOT / "include" / "loki_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(loki_\w+)\s*\(", text,
                       flags=re.M)
    return sorted(set(names))


def test_c_abi_exports_every_declared_symbol():
    import loki_amd

    lib = ctypes.CDLL(str(loki_amd.LIBRARY_PATH))
    names = _declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_c_abi_error_convention():
    """Invalid config -> LOKI_ERR_INVALID (1) and a thread-local message."""
    import loki_amd

    lib = ctypes.CDLL(str(loki_amd.LIBRARY_PATH))
    lib.loki_last_error.restype = ctypes.c_char_p

    class Params(ctypes.Structure):
        _fields_ = [("nsamps", ctypes.c_size_t), ("tsamp", ctypes.c_double),
                    ("nbins", ctypes.c_size_t), ("eta", ctypes.c_double),
                    ("param_limits", ctypes.POINTER(ctypes.c_double)),
                    ("nparams", ctypes.c_size_t), ("ducy_max", ctypes.c_double),
                    ("wtsp", ctypes.c_double), ("use_fourier", ctypes.c_int),
                    ("nthreads", ctypes.c_int), ("max_process_memory_gb", ctypes.c_double),
                    ("octave_scale", ctypes.c_double), ("nbins_max", ctypes.c_size_t),
                    ("nbins_min_lossy_bf", ctypes.c_size_t), ("bseg_brute", ctypes.c_size_t),
                    ("bseg_ffa", ctypes.c_size_t), ("snr_min", ctypes.c_double),
                    ("max_passing_candidates", ctypes.c_size_t),
                    ("prune_poly_order", ctypes.c_size_t), ("p_orb_min", ctypes.c_double),
                    ("m_c_max", ctypes.c_double), ("m_p_min", ctypes.c_double),
                    ("minimum_snap_cells", ctypes.c_double),
                    ("use_conservative_tile", ctypes.c_int)]

    lims = (ctypes.c_double * 2)(1.0, 2.0)
    p = Params(nsamps=1000, tsamp=64e-6, nbins=64, eta=1.0, param_limits=lims, nparams=1,
               ducy_max=0.2, wtsp=1.5, nthreads=1, max_process_memory_gb=8.0,
               nbins_max=1024, nbins_min_lossy_bf=64, bseg_brute=8)
    h = ctypes.c_void_p()
    rc = lib.loki_config_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == 1
    assert b"power of 2" in lib.loki_last_error()
    p.nsamps = 2**16
    rc = lib.loki_config_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == 0 and h.value
    lib.loki_config_destroy(h)


CONFIGS = [
    (2**20, 64, [(0.5, 2.0)], {}),
    (2**23, 32, [(-50.0, 50.0), (100.0, 400.0)], {}),
    (2**22, 32, [(-2.0, 2.0), (-30.0, 30.0), (50.0, 300.0)], {}),
    (2**18, 100, [(1.0, 3.0)], {}),
    (2**17, 32, [(-60000.0, 60000.0), (100.0, 110.0)], {"eta": 2.0, "bseg_brute": 2048}),
]


@pytest.mark.parametrize(("nsamps", "nbins", "lims", "extra"), CONFIGS)
@pytest.mark.parametrize("fourier", [False, True])
def test_plan_matches_oracle(nsamps, nbins, lims, extra, fourier):
    from loki_amd import libloki
    from oracle.plan import Plan, SearchConfig

    eta = extra.get("eta", 1.0)
    kw = {k: v for k, v in extra.items() if k != "eta"}
    cfg = libloki.configs.PulsarSearchConfig(nsamps=nsamps, tsamp=64e-6, nbins=nbins, eta=eta,
                                             param_limits=np.array(lims), use_fourier=fourier,
                                             **kw)
    ocfg = SearchConfig(nsamps, 64e-6, nbins, eta, lims, use_fourier=fourier, **kw)
    plan = (libloki.plans.FFAPlanFourier if fourier else libloki.plans.FFAPlanTime)(cfg)
    oplan = Plan(ocfg, fourier=fourier)
    assert cfg.bseg_brute == ocfg.bseg_brute and cfg.niters_ffa == ocfg.niters_ffa
    assert list(cfg.score_widths) == list(ocfg.score_widths)
    assert cfg.param_names == list(ocfg.param_names)
    assert [list(s) for s in plan.fold_shapes] == oplan.fold_shapes
    assert [list(s) for s in plan.fold_shapes_time] == oplan.fold_shapes_time
    assert list(plan.ncoords) == oplan.ncoords
    assert list(plan.ncoords_offsets) == [x & 0xFFFFFFFF for x in oplan.ncoords_offsets]
    assert list(plan.nsegments) == oplan.nsegments
    assert list(plan.segment_lens) == oplan.segment_lens
    np.testing.assert_array_equal(plan.tsegments, oplan.tsegments)
    for lvl in range(plan.n_levels):
        np.testing.assert_array_equal(plan.dparams[lvl], oplan.dparams[lvl])
        np.testing.assert_array_equal(plan.dparams_actual[lvl], oplan.dparams_act[lvl])
        assert list(plan.param_counts[lvl]) == oplan.param_counts[lvl]
        assert list(plan.param_cart_strides[lvl]) == oplan.param_cart_strides(lvl)
    assert plan.buffer_size == oplan.buffer_size
    assert plan.fold_size == oplan.fold_size and plan.brute_fold_size == oplan.brute_fold_size
    assert plan.coord_size == oplan.coord_size
    assert plan.buffer_memory_usage == pytest.approx(oplan.buffer_memory_usage, rel=1e-6)
    for name, grid in plan.params_dict.items():
        np.testing.assert_array_equal(grid, oplan.params_dict[name])
    np.testing.assert_array_equal(cfg.dparams(1.5), np.array(ocfg.dparams(1.5)))


def test_plan_known_answer_ep_accel():
    """notebooks/loki_ep_accel.ipynb:208-209 through the product's plan class."""
    from loki_amd import libloki

    cfg = libloki.configs.PulsarSearchConfig(
        nsamps=2**25, tsamp=64e-6, nbins=64, eta=1.0,
        param_limits=np.array([[-2500.0, 2500.0], [138.74643398, 151.34310645]]),
        bseg_brute=2**25 // 16384, bseg_ffa=2**25 // 128, ducy_max=0.5, wtsp=1.2)
    plan = libloki.plans.FFAPlanFourier(cfg)
    assert list(plan.fold_shapes[0]) == [16384, 1, 53, 2, 33]
    assert list(plan.fold_shapes[-1]) == [128, 3, 6763, 2, 33]
    assert round(plan.buffer_memory_usage, 2) == 2.55


@pytest.mark.parametrize(("kw", "msg"), [
    (dict(nsamps=1000, bseg_brute=8), "power of 2"),
    (dict(nsamps=2**16, bseg_brute=3), "bseg_brute"),
    (dict(nsamps=2**16, bseg_brute=2**16), "bseg_brute must be less than nsamps"),
    (dict(nsamps=2**16, bseg_brute=1024, bseg_ffa=512), "bseg_ffa"),
    (dict(nsamps=2**16, bseg_brute=1024, nbins=2048), "nbins_max"),
])
def test_config_errors_raise_runtime_error(kw, msg):
    from loki_amd import libloki

    args = dict(tsamp=64e-6, nbins=64, eta=1.0, param_limits=np.array([[1.0, 2.0]]))
    args.update(kw)
    with pytest.raises(RuntimeError, match=msg):
        libloki.configs.PulsarSearchConfig(**args)


def test_param_limits_shape_error():
    from loki_amd import libloki

    with pytest.raises((ValueError, RuntimeError)):
        libloki.configs.PulsarSearchConfig(nsamps=2**16, tsamp=64e-6, nbins=64, eta=1.0,
                                           param_limits=np.array([1.0, 2.0]))


def test_generate_box_width_trials_binding_default():
    """loki_python.cpp:78: the Python binding defaults to ducy_max = 0.3."""
    from loki_amd import libloki

    assert list(libloki.scores.generate_box_width_trials(64)) == [1, 2, 3, 4, 6, 9, 13, 19]
    assert list(libloki.scores.generate_box_width_trials(64, 0.2, 1.5)) == [1, 2, 3, 4, 6, 9]


def test_resolve_coordinates_needs_matching_kind():
    from loki_amd import libloki

    cfg = libloki.configs.PulsarSearchConfig(nsamps=2**16, tsamp=64e-6, nbins=64, eta=1.0,
                                             param_limits=np.array([[2.0, 8.0]]),
                                             use_fourier=False)
    with pytest.raises(RuntimeError, match="nparams>=2"):
        libloki.plans.FFAPlanTime(cfg).resolve_coordinates()


def test_libloki_ffa_namespace_matches_the_cpu_module():
    """search.py:7 imports ``loki.libloki.ffa.compute_ffa_scores``; the drop-in module has the
    names of src/loki_python.cpp:518-592 with the same argument names and defaults, importable
    by dotted name."""
    import inspect

    from loki_amd.libloki.ffa import (FFAFourier, FFAFreqSweep, FFATime, compute_ffa_fourier,
                                      compute_ffa_fourier_return_to_time, compute_ffa_scores,
                                      compute_ffa_time)

    for fn in (compute_ffa_time, compute_ffa_fourier, compute_ffa_fourier_return_to_time,
               compute_ffa_scores):
        sig = inspect.signature(fn)
        assert list(sig.parameters) == ["ts_e", "ts_v", "cfg", "quiet", "show_progress"]
        assert sig.parameters["quiet"].default is False
        assert sig.parameters["show_progress"].default is False
    for cls in (FFATime, FFAFourier, FFAFreqSweep):
        sig = inspect.signature(cls.__init__)
        assert list(sig.parameters) == ["self", "cfg", "show_progress"]
        assert sig.parameters["show_progress"].default is True
    assert list(inspect.signature(FFATime.execute).parameters) == ["self", "ts_e", "ts_v", "fold"]
    assert list(inspect.signature(FFAFreqSweep.execute).parameters) == [
        "self", "ts_e", "ts_v", "outdir", "file_prefix"]
    from loki_amd.libloki.configs import PulsarSearchConfig  # noqa: F401
    from loki_amd.libloki.plans import FFAPlanTime  # noqa: F401
    from loki_amd.libloki.scores import generate_box_width_trials  # noqa: F401
    from loki_amd.libculoki.ffa import FFACUDA, compute_ffa_scores_cuda  # noqa: F401
