"""Exact scoring's normalisation p = e * (1 / sqrt(v)) (score.cpp:139-141) runs, inside fixed
ranges, on short correctly rounded sequences instead of the IEEE square root and division.
These tests pin them, through the C-ABI self-test entry points:

* exhaustively: for every float in the root's range [2^-96, FLT_MAX] and the reciprocal's range
  [2^-48, 2^64) the short sequence equals the device's IEEE sqrtf / 1.0f / s bit for bit;
* the device's IEEE operations themselves equal the host's (numpy float32, IEEE) on random bit
  patterns of every class, including zeros, subnormals, negatives, inf and NaN, which the exact
  kernels route to the IEEE operations.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT_LO, ROOT_HI = 0x0F800000, 0x7F800000    # [2^-96, inf)
RCP_LO, RCP_HI = 0x27800000, 0x5F800000      # [2^-48, 2^64)


def _lib():
    import loki_amd

    lib = ctypes.CDLL(str(loki_amd.LIBRARY_PATH))
    lib.loki_selftest_exact_norm_sweep.argtypes = [
        ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
        ctypes.POINTER(ctypes.c_uint32)]
    lib.loki_selftest_exact_norm_values.argtypes = [
        ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.loki_last_error.restype = ctypes.c_char_p
    return lib


def test_short_sqrt_and_reciprocal_equal_ieee_for_every_float():
    lib = _lib()
    counts = (ctypes.c_uint64 * 4)()
    first = (ctypes.c_uint32 * 2)()
    # every non-negative pattern below +inf: both ranges lie inside
    rc = lib.loki_selftest_exact_norm_sweep(0, 0, 0x7F800000, counts, first)
    assert rc == 0, lib.loki_last_error().decode()
    assert counts[2] == ROOT_HI - ROOT_LO      # every float of the root's range was checked
    assert counts[3] == RCP_HI - RCP_LO        # ... and of the reciprocal's
    assert counts[0] == 0, f"sqrt mismatches {counts[0]}, first at {first[0]:#010x}"
    assert counts[1] == 0, f"reciprocal mismatches {counts[1]}, first at {first[1]:#010x}"


def test_device_ieee_sqrt_and_division_equal_host():
    lib = _lib()
    rng = np.random.default_rng(2024)
    bits = rng.integers(0, 2**32, size=2**21, dtype=np.uint64).astype(np.uint32)
    edges = np.array([0x00000000, 0x80000000, 0x00000001, 0x007FFFFF, 0x00800000, 0x0F7FFFFF,
                      0x0F800000, 0x27800000, 0x277FFFFF, 0x5F7FFFFF, 0x5F800000, 0x7F7FFFFF,
                      0x7F800000, 0xFF800000, 0x7FC00000, 0xBF800000, 0x3F800000],
                     np.uint32)
    bits = np.concatenate([bits, edges])
    out = np.empty((bits.size, 4), np.float32)
    rc = lib.loki_selftest_exact_norm_values(0, bits.ctypes.data, bits.size, out.ctypes.data)
    assert rc == 0, lib.loki_last_error().decode()
    x = bits.view(np.float32)
    with np.errstate(all="ignore"):
        want_sqrt = np.sqrt(x)
        want_rcp = np.float32(1.0) / x
    np.testing.assert_array_equal(out[:, 0], want_sqrt)   # NaN positions equal too
    np.testing.assert_array_equal(out[:, 1], want_rcp)
    in_root = (bits >= ROOT_LO) & (bits < ROOT_HI)
    in_rcp = (bits >= RCP_LO) & (bits < RCP_HI)
    assert in_root.sum() > 10**5 and in_rcp.sum() > 10**5
    np.testing.assert_array_equal(out[in_root, 2], want_sqrt[in_root])
    np.testing.assert_array_equal(out[in_rcp, 3], want_rcp[in_rcp])
