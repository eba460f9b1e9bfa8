"""ctypes wrappers around oracle/liboracle.so and whole-pipeline drivers.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  The drivers mirror the
level loop of ffa.cpp:242-360 and the compute_* entry points of
ffa.cpp:405-466 so tests can compare the HIP product end to end.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from .plan import FFA_COORD_DTYPE, FFA_COORD_FREQ_DTYPE, Plan, SearchConfig

_HERE = Path(__file__).resolve().parent
_LIB = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_sz = C.c_size_t


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        # LOKI_ORACLE_LIB: another build of the same source (the sanitizer test's
        # -fsanitize=address,undefined build)
        alt = os.environ.get("LOKI_ORACLE_LIB")
        path = Path(alt) if alt else _HERE / "liboracle.so"
        if not path.exists():
            if alt:
                raise FileNotFoundError(alt)
            build()
        L = C.CDLL(str(path))
        L.orc_phase_idx.restype = C.c_float
        L.orc_phase_idx.argtypes = [C.c_double, C.c_double, _sz, C.c_double]
        L.orc_brute_fold_time.argtypes = [_f32p, _f32p, _sz, _f64p, _sz, _sz, _sz, C.c_double,
                                          C.c_double, _f32p, C.c_int]
        L.orc_brute_fold_time_f64.argtypes = [_f32p, _f32p, _sz, _f64p, _sz, _sz, _sz,
                                              C.c_double, C.c_double, _f64p, C.c_int]
        for name, p in (("orc_ffa_iter_freq", _f32p), ("orc_ffa_iter_poly", _f32p),
                        ("orc_ffa_iter_freq_f64", _f64p), ("orc_ffa_iter_poly_f64", _f64p),
                        ("orc_ffa_complex_iter_freq", _f32p), ("orc_ffa_complex_iter_poly", _f32p)):
            getattr(L, name).argtypes = [p, p, C.c_void_p, _sz, _sz, _sz, _sz, C.c_int]
        L.orc_brute_fold_fourier.argtypes = [_f32p, _f32p, _sz, _f64p, _sz, _sz, _sz, C.c_double,
                                             C.c_double, _f32p, C.c_int]
        L.orc_irfft.argtypes = [_f32p, _f32p, _sz, _sz, C.c_int]
        L.orc_circular_prefix_sum.argtypes = [_f32p, _f32p, _sz, _sz]
        L.orc_diff_max.restype = C.c_float
        L.orc_diff_max.argtypes = [_f32p, _f32p, _sz]
        L.orc_snr_boxcar_3d.argtypes = [_f32p, _sz, _sz, _u64p, _sz, _f32p, C.c_int, C.c_int]
        L.orc_snr_boxcar_3d_f64.argtypes = [_f64p, _sz, _sz, _u64p, _sz, _f64p, C.c_int]
        L.orc_round_storage.argtypes = [_f32p, _sz, C.c_float, C.c_int, C.c_int]
        L.orc_round_bfp16.argtypes = [_f32p, _sz, _sz, C.c_int]
        L.orc_bfp16_exponent.restype = C.c_int
        L.orc_bfp16_exponent.argtypes = [C.c_float]
        _LIB = L
    return _LIB


def _nthreads(n: int | None) -> int:
    return int(n) if n else (os.cpu_count() or 1)


# ---------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------
def brute_fold_time(ts_e, ts_v, freqs, segment_len, nbins, tsamp, t_ref=0.0, nthreads=None,
                    f64=False):
    ts_e = np.ascontiguousarray(ts_e, np.float32)
    ts_v = np.ascontiguousarray(ts_v, np.float32)
    freqs = np.ascontiguousarray(freqs, np.float64)
    nseg = ts_e.size // segment_len
    if f64:
        out = np.zeros(nseg * freqs.size * 2 * nbins, np.float64)
        lib().orc_brute_fold_time_f64(ts_e, ts_v, ts_e.size, freqs, freqs.size, segment_len, nbins,
                                      tsamp, t_ref, out, _nthreads(nthreads))
    else:
        out = np.zeros(nseg * freqs.size * 2 * nbins, np.float32)
        lib().orc_brute_fold_time(ts_e, ts_v, ts_e.size, freqs, freqs.size, segment_len, nbins,
                                  tsamp, t_ref, out, _nthreads(nthreads))
    return out


def brute_fold_fourier(ts_e, ts_v, freqs, segment_len, nbins, tsamp, t_ref=0.0, nthreads=None):
    ts_e = np.ascontiguousarray(ts_e, np.float32)
    ts_v = np.ascontiguousarray(ts_v, np.float32)
    freqs = np.ascontiguousarray(freqs, np.float64)
    nseg = ts_e.size // segment_len
    nbf = nbins // 2 + 1
    out = np.zeros(nseg * freqs.size * 2 * nbf * 2, np.float32)
    lib().orc_brute_fold_fourier(ts_e, ts_v, ts_e.size, freqs, freqs.size, segment_len, nbins,
                                 tsamp, t_ref, out, _nthreads(nthreads))
    return out.view(np.complex64)


def ffa_iter(fold_in, coords, ncur, nprev, nseg, nbins, nthreads=None, fourier=False):
    """One merge level; dispatch on coord dtype (freq-only vs poly) and domain."""
    coords = np.ascontiguousarray(coords)
    freq_only = coords.dtype == FFA_COORD_FREQ_DTYPE
    if fourier:
        nbf = nbins // 2 + 1
        fin = np.ascontiguousarray(fold_in).view(np.float32)
        out = np.empty(nseg * ncur * 2 * nbf * 2, np.float32)
        fn = lib().orc_ffa_complex_iter_freq if freq_only else lib().orc_ffa_complex_iter_poly
        fn(fin, out, coords.ctypes.data, ncur, nprev, nseg, nbins, _nthreads(nthreads))
        return out.view(np.complex64)
    f64 = fold_in.dtype == np.float64
    out = np.empty(nseg * ncur * 2 * nbins, np.float64 if f64 else np.float32)
    if f64:
        fn = lib().orc_ffa_iter_freq_f64 if freq_only else lib().orc_ffa_iter_poly_f64
    else:
        fn = lib().orc_ffa_iter_freq if freq_only else lib().orc_ffa_iter_poly
    fn(np.ascontiguousarray(fold_in), out, coords.ctypes.data, ncur, nprev, nseg, nbins,
       _nthreads(nthreads))
    return out


def irfft(fold_c, nbins, nthreads=None):
    fin = np.ascontiguousarray(fold_c, np.complex64).view(np.float32)
    nbf = nbins // 2 + 1
    nfft = fin.size // (2 * nbf)
    out = np.empty(nfft * nbins, np.float32)
    lib().orc_irfft(fin, out, nfft, nbins, _nthreads(nthreads))
    return out


def rfft(fold_t, nbins):
    """Forward real DFT of rows of nbins floats (FP64 numpy, rounded to complex64): the
    unnormalised R2C transform of fft_cuda.cu:282-355 / fft.cpp:375-500."""
    x = np.ascontiguousarray(fold_t, np.float32).reshape(-1, nbins).astype(np.float64)
    return np.fft.rfft(x, axis=1).astype(np.complex64).reshape(-1)


def circular_prefix_sum(x, nsum):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(nsum, np.float32)
    lib().orc_circular_prefix_sum(x, out, x.size, nsum)
    return out


def diff_max(x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return float(lib().orc_diff_max(x, y, min(x.size, y.size)))


def snr_boxcar_3d(folds, widths, nbins=None, find_max=False, nthreads=None):
    folds = np.ascontiguousarray(folds, np.float32)
    widths = np.ascontiguousarray(widths, np.uint64)
    if nbins is None:
        nbins = folds.shape[-1]
    nprof = folds.size // (2 * nbins)
    out = np.empty(nprof if find_max else nprof * widths.size, np.float32)
    lib().orc_snr_boxcar_3d(folds.reshape(-1), nprof, nbins, widths, widths.size, out,
                            int(find_max), _nthreads(nthreads))
    return out if find_max else out.reshape(nprof, widths.size)


def snr_boxcar_3d_f64(folds, widths, nbins, nthreads=None):
    folds = np.ascontiguousarray(folds, np.float64)
    widths = np.ascontiguousarray(widths, np.uint64)
    nprof = folds.size // (2 * nbins)
    out = np.empty(nprof * widths.size, np.float64)
    lib().orc_snr_boxcar_3d_f64(folds.reshape(-1), nprof, nbins, widths, widths.size, out,
                                _nthreads(nthreads))
    return out.reshape(nprof, widths.size)


# ---------------------------------------------------------------------------
# drivers (ffa.cpp:242-360, 405-466)
# ---------------------------------------------------------------------------
def round_storage(x, storage: str):
    """Round FP32 values to the stored precision (round to nearest even) and return them as
    the exactly equal FP32 values: 'f32' (unchanged), 'f16' (numpy's IEEE binary16
    conversion) or 'bf16' (the upper 16 bits of the FP32 pattern, rounded to nearest even)."""
    x = np.ascontiguousarray(x, np.float32)
    if storage == "f32":
        return x
    if storage == "f16":
        with np.errstate(over="ignore"):  # beyond 65504 rounds to inf, as on the GPU
            return x.astype(np.float16).astype(np.float32)
    if storage == "bf16":
        u = x.view(np.uint32).astype(np.uint64)
        u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
        return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    raise ValueError(f"unknown storage {storage!r}")


_STORAGE_CODE = {"f32": 0, "f16": 1, "bf16": 2}
STORAGES = ("f32", "f16", "bf16", "bfp16")


def bfp16_exponent(mx):
    """Shared exponent E of a block floating-point row whose largest magnitude is mx (float32):
    max(ex(mx) - 14, -126), ex the unbiased exponent field (-127 for zero and subnormals)."""
    bits = np.asarray(mx, np.float32).view(np.uint32).astype(np.int64)
    return np.maximum(((bits >> 23) & 0xFF) - 127 - 14, -126)


def round_bfp16(x, nbins: int):
    """Block floating-point storage of rows of nbins float32 values (numpy): each row becomes
    int16 mantissas m times 2^E with E = bfp16_exponent(max |row|), m = rint(x 2^-E) (ties to
    even) clamped to +-32767; returns the decoded float32 values m 2^E (exact)."""
    x = np.ascontiguousarray(x, np.float32)
    rows = x.reshape(-1, nbins)
    e = bfp16_exponent(np.abs(rows).max(axis=1))[:, None]
    inv = np.ldexp(np.float32(1.0), -e).astype(np.float32)
    m = np.clip(np.rint(rows * inv), -32767, 32767).astype(np.float32)
    return (m * np.ldexp(np.float32(1.0), e).astype(np.float32)).reshape(x.shape)


def scale_round_storage(x, scale: float, storage: str, nthreads=None):
    """round_storage(x * scale, storage) in place on a float32 array, in C with OpenMP (the
    same rounding; used on whole fold levels, where numpy's temporaries would not fit)."""
    if x.dtype != np.float32 or not x.flags.c_contiguous:
        raise ValueError("scale_round_storage needs a C-contiguous float32 array")
    flat = x.reshape(-1)
    if storage == "bfp16":
        raise ValueError("bfp16 rounds whole rows: use round_bfp16_rows")
    lib().orc_round_storage(flat, flat.size, float(scale), _STORAGE_CODE[storage],
                            _nthreads(nthreads))
    return x


def round_bfp16_rows(x, nbins: int, nthreads=None):
    """round_bfp16 in place on a float32 array of whole rows, in C with OpenMP."""
    if x.dtype != np.float32 or not x.flags.c_contiguous or x.size % nbins:
        raise ValueError("round_bfp16_rows needs a C-contiguous float32 array of whole rows")
    flat = x.reshape(-1)
    lib().orc_round_bfp16(flat, flat.size // nbins, nbins, _nthreads(nthreads))
    return x


def storage_k0(cfg: SearchConfig, plan: Plan) -> int:
    """Level-0 scale exponent of reduced-precision storage: floor(log2(seglen // nbins))."""
    per, k0 = plan.segment_lens[0] // cfg.nbins, 0
    while per > 1:
        per >>= 1
        k0 += 1
    return k0


def compute_ffa(ts_e, ts_v, cfg: SearchConfig, fourier: bool = False, nthreads=None,
                f64: bool = False, return_levels: bool = False, storage: str = "f32"):
    """Run brute fold + all merge levels; returns (final fold, plan[, levels]).

    storage 'f16' / 'bf16' (time domain): levels 0..L-2 are stored at that precision, each
    value rounded once from its FP32 sum at the exact scale 2^-(k0 + l) (merge = (t + h) * 0.5);
    the last level is returned in FP32 at its true scale, (t + h) * 2^(k0 + L - 2).
    storage 'bfp16': level 0 (the brute fold) stays FP32, levels 1..L-2 are stored as block
    floating point (round_bfp16 of each component row of the FP32 sum t + h, no scaling) and
    the last level is the FP32 sum.
    """
    if storage != "f32":
        if fourier or f64:
            raise ValueError("reduced-precision storage applies to the FP32 time-domain FFA")
        return _compute_ffa_storage(ts_e, ts_v, cfg, nthreads, storage)
    plan = Plan(cfg, fourier=fourier)
    coords = plan.resolve_coordinates()
    freqs = plan.brute_fold_freqs()
    t_ref = plan.brute_fold_tref()
    seglen = plan.segment_lens[0]
    if fourier and cfg.nbins > cfg.nbins_min_lossy_bf:
        # lossy init (ffa.cpp:180-192, 207-234): time brute fold, then rfft
        fold_t = brute_fold_time(ts_e, ts_v, freqs, seglen, cfg.nbins, cfg.tsamp, t_ref, nthreads)
        fold = rfft(fold_t, cfg.nbins)
    elif fourier:
        fold = brute_fold_fourier(ts_e, ts_v, freqs, seglen, cfg.nbins, cfg.tsamp, t_ref, nthreads)
    else:
        fold = brute_fold_time(ts_e, ts_v, freqs, seglen, cfg.nbins, cfg.tsamp, t_ref, nthreads,
                               f64=f64)
    levels = [fold]
    for lvl in range(1, plan.n_levels):
        fold = ffa_iter(fold, coords[lvl], plan.ncoords[lvl], plan.ncoords[lvl - 1],
                        plan.nsegments[lvl], cfg.nbins, nthreads, fourier=fourier)
        if return_levels:
            levels.append(fold)
    if return_levels:
        return fold, plan, levels
    return fold, plan


def _compute_ffa_storage(ts_e, ts_v, cfg: SearchConfig, nthreads, storage: str):
    plan = Plan(cfg)
    coords = plan.resolve_coordinates()
    freqs = plan.brute_fold_freqs()
    t_ref = plan.brute_fold_tref()
    if storage == "bfp16":
        fold = brute_fold_time(ts_e, ts_v, freqs, plan.segment_lens[0], cfg.nbins, cfg.tsamp,
                               t_ref, nthreads)
        last = plan.n_levels - 1
        for lvl in range(1, plan.n_levels):
            out = ffa_iter(fold, coords[lvl], plan.ncoords[lvl], plan.ncoords[lvl - 1],
                           plan.nsegments[lvl], cfg.nbins, nthreads)
            del fold
            fold = round_bfp16_rows(out, cfg.nbins, nthreads) if lvl < last else out
        return fold, plan
    k0 = storage_k0(cfg, plan)
    fold = brute_fold_time(ts_e, ts_v, freqs, plan.segment_lens[0], cfg.nbins, cfg.tsamp, t_ref,
                           nthreads)
    last = plan.n_levels - 1
    if last == 0:
        return fold, plan
    fold = scale_round_storage(fold, 2.0 ** -k0, storage, nthreads)
    for lvl in range(1, plan.n_levels):
        out = ffa_iter(fold, coords[lvl], plan.ncoords[lvl], plan.ncoords[lvl - 1],
                       plan.nsegments[lvl], cfg.nbins, nthreads)
        del fold
        if lvl < last:
            fold = scale_round_storage(out, 0.5, storage, nthreads)
        else:
            fold = scale_round_storage(out, 2.0 ** (k0 + lvl - 1), "f32", nthreads)
    return fold, plan


def compute_ffa_scores(ts_e, ts_v, cfg: SearchConfig, widths=None, nthreads=None,
                       storage: str = "f32"):
    """ffa.cpp:444-466: FFA (time, or Fourier then irfft) + snr_boxcar_3d."""
    if cfg.use_fourier:
        fold_c, plan = compute_ffa(ts_e, ts_v, cfg, fourier=True, nthreads=nthreads)
        fold = irfft(fold_c, cfg.nbins, nthreads)
    else:
        fold, plan = compute_ffa(ts_e, ts_v, cfg, fourier=False, nthreads=nthreads,
                                 storage=storage)
    if plan.nsegments[-1] != 1:
        raise ValueError("compute_ffa_scores: nsegments must be 1 for scores")
    w = cfg.score_widths if widths is None else np.asarray(widths, np.uint64)
    return snr_boxcar_3d(fold, w, cfg.nbins, nthreads=nthreads).reshape(-1), plan


__all__ = [
    "FFA_COORD_DTYPE", "FFA_COORD_FREQ_DTYPE", "Plan", "SearchConfig", "brute_fold_fourier",
    "STORAGES", "bfp16_exponent", "brute_fold_time", "build", "circular_prefix_sum",
    "compute_ffa", "compute_ffa_scores", "round_bfp16", "round_bfp16_rows",
    "diff_max", "ffa_iter", "irfft", "lib", "rfft", "round_storage", "snr_boxcar_3d",
    "scale_round_storage", "snr_boxcar_3d_f64", "storage_k0",
]
