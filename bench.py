#!/usr/bin/env python3
"""Benchmark of the FFA hot path on one MI355X.

Workload (default): Taylor FFA over (accel, freq) + boxcar S/N, FP32, time domain --
2^23 float32 samples at tsamp = 64 us, N(0,1) noise from a fixed seed with an injected binary
pulsar (f0 = 200 Hz, accel = +20 m/s^2, Gaussian pulse FWHM 4 %, S/N 15); grid f 100-400 Hz,
accel -50..+50 m/s^2 at the FFA-native spacing (eta = 1), nbins = 32, boxcar widths
{1,2,3,4,6,8,12,16,24} (the list {1,...,32} minus widths >= nbins).

One step = shift tables for every level + brute fold + all merge levels + S/N of every final
trial, on inputs already resident in HBM.  value = final-level parameter trials per second.

Prints ONE JSON line.  ``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from collections import defaultdict
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "FFA param trials/s + % HBM roofline; FP16-store vs FP32 at |\u0394S/N|<1e-3"
HBM_PEAK_GBPS = 8000.0
MFMA_F32_PEAK_TFLOPS = 157.3
# VALU issue: 256 CUs x 4 SIMDs at 2.4 GHz, a wave64 instruction per SIMD every 4 cycles, and
# up to two per 4 cycles when the SIMD dual-issues (SQ_ACTIVE_INST_VALU2 counts those
# quad-cycles) -> 78.6e12 lane-instructions/s, the published 157.3 TFLOP/s FP32 vector peak
# counted as FMAs (2 flop each).
SIMDS = 1024
CLOCK_HZ = 2.4e9
VALU_LANE_INST_PEAK_T = SIMDS * 16 * 2 * CLOCK_HZ / 1e12
# LDS: 128 bytes per clock per CU (the CDNA3 figure; if gfx950 reads faster, the fraction
# below overstates the LDS utilisation) -> 78.6 TB/s over 256 CUs.
LDS_PEAK_TBPS = 256 * 128 * CLOCK_HZ / 1e12
WIDTHS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32)

WORKLOADS = {
    "taylor_ffa_accel_freq": dict(
        nsamps=2**23, tsamp=64e-6, nbins=32, eta=1.0,
        param_limits=[(-50.0, 50.0), (100.0, 400.0)],
        signal=dict(f0=200.0, accel=20.0, jerk=0.0, fwhm=0.04, snr=15.0), seed=20231,
        # CPU baseline on the whole workload: ~10 s on 16 threads, ~75 GB of host memory
        # (two fold levels, the shift tables and the scores).
        cpu_sample=dict(param_limits=[(-50.0, 50.0), (100.0, 400.0)], bseg_brute=1024),
        # S/N error of each fold storage against an FP64 oracle fold: the full accel range
        # over a 2 Hz band around the injection
        storage_sample=dict(param_limits=[(-50.0, 50.0), (199.0, 201.0)]),
    ),
    # Fourier-domain FFA over (fddot, fdot, f) expressed as (jerk, accel, freq) at f_min = 50 Hz:
    # accel = c * fdot / f_min, jerk = c * fddot / f_min for |fdot| <= 1e-8 Hz/s,
    # |fddot| <= 1e-14 Hz/s^2; complex64 folds with nbins 64 (33 harmonics).
    "fourier_ffa_fdot_fddot": dict(
        nsamps=2**22, tsamp=64e-6, nbins=64, eta=1.0, fourier=True,
        param_limits=[(-299792458.0 * 1e-14 / 50.0, 299792458.0 * 1e-14 / 50.0),
                      (-299792458.0 * 1e-8 / 50.0, 299792458.0 * 1e-8 / 50.0), (50.0, 500.0)],
        signal=dict(f0=251.3, accel=0.0, jerk=0.0, fwhm=0.05, snr=15.0), seed=5150,
        # 64 injections (f in [50, 500] Hz, |fdot| <= 1e-8 Hz/s, |fddot| <= 1e-14 Hz/s^2,
        # S/N 10-25); the 4096 strongest grid peaks are refined by gradient ascent of the
        # fixed-width S/N in (f, fdot, fddot) (loki_amd.refine).
        injections=dict(n=64, seed=5151, f=(50.0, 500.0), fdot=1e-8, fddot=1e-14,
                        snr=(10.0, 25.0), fwhm=0.05),
        refine=dict(ncand=4096, iters=10),
        cpu_sample=dict(param_limits=[(-299792458.0 * 1e-14 / 50.0, 299792458.0 * 1e-14 / 50.0),
                                      (-299792458.0 * 1e-8 / 50.0, 299792458.0 * 1e-8 / 50.0),
                                      (496.0, 500.0)], bseg_brute=2048),
    ),
    # Jerk search with reduced-precision fold storage: levels 0..L-2 stored as f32, f16 or bf16
    # with every add in FP32; the S/N error of each storage against an FP64 oracle fold is
    # measured on a frequency sub-band around the injection (full accel/jerk range).
    "jerk_ffa_fold_storage": dict(
        nsamps=2**22, tsamp=64e-6, nbins=32, eta=1.0,
        param_limits=[(-2.0, 2.0), (-30.0, 30.0), (50.0, 300.0)],
        signal=dict(f0=150.0, accel=10.0, jerk=0.5, fwhm=0.05, snr=12.0), seed=4242,
        # CPU baseline on the whole workload: ~6 s on 16 threads, ~70 GB of host memory
        cpu_sample=dict(param_limits=[(-2.0, 2.0), (-30.0, 30.0), (50.0, 300.0)],
                        bseg_brute=None),
        storage_sample=dict(param_limits=[(-2.0, 2.0), (-30.0, 30.0), (149.0, 151.0)]),
    ),
    # 256 independent period searches of uneven length (2^18..2^22 samples), nbins in
    # {32, 64, 128}, period ranges [P/2, 2P] (loki_amd.synth.make_batch_specs), run as one
    # batch: grouped launches per stage captured in a HIP graph, autotuned tile schedules.
    "batched_uneven": dict(
        batch=True, nseries=256, seed=7000, tsamp=64e-6, nsamps=2**22, nbins=128, eta=1.0,
        cpu_sample=dict(nseries=12),
    ),
    # Frequency sweep over four octaves (25-400 Hz) with an accel range: the region planner's
    # chunks (nbins 32..256, eta 1..8) each run FFA + score-and-filter at S/N > 7 into one
    # candidate list, written to a result archive (FFAFreqSweepCUDA).
    "freq_sweep_accel": dict(
        sweep=True, nsamps=2**22, tsamp=64e-6, nbins=32, eta=1.0,
        param_limits=[(-30.0, 30.0), (25.0, 400.0)], snr_min=7.0, max_process_memory_gb=64.0,
        injections=[dict(f0=f, fdot=0.0, fddot=0.0, fwhm=0.05, snr=15.0)
                    for f in (31.7, 77.3, 151.9, 333.3)], seed=8080,
    ),
    "ffa_freq_only": dict(
        nsamps=2**20, tsamp=64e-6, nbins=64, eta=1.0, param_limits=[(0.5, 2.0)],
        signal=dict(f0=1.0, accel=0.0, jerk=0.0, fwhm=0.05, snr=20.0), seed=101,
        cpu_sample=dict(param_limits=[(0.5, 2.0)], bseg_brute=None),
    ),
}


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="taylor_ffa_accel_freq", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fast-norm", action="store_true",
                    help="hardware rsqrt profile normalisation (S/N within 1e-4, not bit-exact)")
    ap.add_argument("--unfused", action="store_true",
                    help="store the final fold and score it in a separate kernel")
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16", "bfp16"],
                    help="fold storage precision of levels 0..L-2 (adds stay FP32; bfp16: "
                         "block floating point, level 0 in FP32)")
    ap.add_argument("--no-storage-compare", action="store_true",
                    help="skip timing the FP16 / BF16 storage variants beside the main run")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: every CPU this process may use (affinity, capped by OMP_NUM_THREADS)")
    ap.add_argument("--fuse-depth", type=int, default=2, choices=[1, 2],
                    help="levels fused into the scoring kernel (2: the parent level of the last "
                         "is merged on chip too)")
    ap.add_argument("--no-score-filter", action="store_true",
                    help="skip the score-and-filter timing")
    ap.add_argument("--filter-threshold", type=float, default=6.0,
                    help="S/N threshold of the score-and-filter timing")
    ap.add_argument("--dft", default="f32", choices=["f32", "bf16x3"],
                    help="Fourier brute fold on FP32 matrix cores or split-bf16 ones")
    ap.add_argument("--no-seeds", action="store_true",
                    help="skip the pruning-seed kernel timing (time-domain workloads)")
    ap.add_argument("--no-refine", action="store_true",
                    help="Fourier workload: skip the gradient-ascent refinement leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (3 times): the "
                         "reference kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def calibrate_copy(nbytes: int) -> int:
    import torch

    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4, "launches": 3}), flush=True)
    return 0


def alg_bytes(plan, nbins, nw, kind, level, nsamps, fourier=False):
    """Algorithmic bytes of one launch: FP32 time folds (4 B x nbins per row) or complex64
    Fourier folds (8 B x (nbins/2+1) per row); 16-byte poly / 8-byte freq coords."""
    npar = int(plan.n_params)
    rec = 16 if npar > 1 else 8
    nc = [int(x) for x in plan.ncoords]
    ns = [int(x) for x in plan.nsegments]
    row = 2 * (nbins // 2 + 1) * 8 if fourier else 2 * nbins * 4
    if kind == "merge":
        return ns[level] * nc[level] * row * 3 + nc[level] * rec
    if kind == "resolve":
        return nc[level] * rec
    if kind == "brute_fold":
        return 2 * nsamps * 4 + ns[0] * nc[0] * row
    if kind == "transform":  # final Fourier fold -> time fold
        return nc[-1] * row + nc[-1] * 2 * nbins * 4
    if kind == "score":
        return nc[-1] * 2 * nbins * 4 + nc[-1] * nw * 4
    if kind == "merge_score":  # fused last level: same algorithmic bytes as merge + score
        return (alg_bytes(plan, nbins, nw, "merge", level, nsamps)
                + alg_bytes(plan, nbins, nw, "score", level, nsamps))
    raise ValueError(kind)


def dft_flops(plan, nbins, nsamps):
    """Real multiply-adds x 2 of the Fourier brute fold as a matrix product: rows (segment,
    component) x samples x (re, im) of harmonics 1..nbins/2, per level-0 frequency."""
    nh = nbins // 2
    return 2.0 * (2 * nsamps) * (2 * nh) * int(plan.param_counts[0][-1])


def pipeline_bytes(plan, nbins, nw, nsamps, fourier=False):
    tot = alg_bytes(plan, nbins, nw, "brute_fold", 0, nsamps, fourier)
    if fourier:
        tot += alg_bytes(plan, nbins, nw, "transform", 0, nsamps, fourier)
    for lvl in range(1, int(plan.n_levels)):
        tot += alg_bytes(plan, nbins, nw, "merge", lvl, nsamps, fourier)
    tot += alg_bytes(plan, nbins, nw, "score", 0, nsamps, fourier)
    return tot


def load_traffic(workload: str, kernel: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/hbm_traffic.json)."""
    path = ROOT / "profiles" / "hbm_traffic.json"
    if not path.exists():
        return None
    try:
        data = json.loads(path.read_text())
        return data.get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None


def cpu_share():
    """Threads of the CPU baseline: every CPU this process may run on (its affinity mask),
    capped by OMP_NUM_THREADS when the environment sets it (the host's share for this GPU);
    plus the host facts the baseline is reported with."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    n = min(aff, int(omp)) if omp.isdigit() and int(omp) > 0 else aff
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, {"host_cpus": os.cpu_count(), "affinity_cpus": aff,
               "omp_num_threads": omp or None, "cpu_model": model}


def load_counters(workload: str, kernel: str):
    """Per-launch SQ counters of a kernel measured with rocprofv3 --pmc
    (profiles/valu_counters.json)."""
    path = ROOT / "profiles" / "valu_counters.json"
    if not path.exists():
        return None
    try:
        return json.loads(path.read_text()).get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None


def cpu_baseline(wl: dict, ts_e, ts_v, widths, threads: int) -> dict:
    """The CPU oracle (C/OpenMP restatement of the reference CPU FFA) on a bounded sample."""
    from oracle import oracle as orc
    from oracle.plan import Plan, SearchConfig

    smp = wl["cpu_sample"]
    fourier = bool(wl.get("fourier", False))
    ocfg = SearchConfig(wl["nsamps"], wl["tsamp"], wl["nbins"], wl["eta"], smp["param_limits"],
                        use_fourier=fourier, bseg_brute=smp["bseg_brute"])
    nc = Plan(ocfg, fourier=fourier).ncoords[-1]
    orc.lib()
    t0 = time.perf_counter()
    orc.compute_ffa_scores(ts_e, ts_v, ocfg, widths=widths, nthreads=threads)
    dt = time.perf_counter() - t0
    lims = ", ".join(f"[{a:g}, {b:g}]" for a, b in smp["param_limits"])
    note = (" (FP64 DFT brute fold and FP64 phase rotations: slower than an FP32 CPU path)"
            if fourier else "")
    _, host = cpu_share()
    return {
        "value": nc / dt, "unit": "param trials/s", "cores": threads, "kind": "port",
        "sample": f"same series and grid rules, param_limits {lims}, bseg_brute "
                  f"{ocfg.bseg_brute}: {nc} final trials x {len(widths)} widths in {dt:.2f} s"
                  + note,
        **host,
    }


# FP32 vector peak: the VALU issue peak above with an FMA (2 flop) per lane-instruction;
# 157.3 TFLOP/s, the published figure (dual-issued non-packed FMAs reach it as packed ones do)
VALU_F32_PEAK_TFLOPS = 2.0 * VALU_LANE_INST_PEAK_T
REFINE_FLOPS_PER_TERM = 12.0  # three-term power recurrence (2 FMA) + 4 accumulating FMAs


def grid_peaks(best, ncand):
    """Indices of the ncand strongest local maxima (>= both neighbours) of a 1-D score row."""
    left = np.concatenate([[-np.inf], best[:-1]])
    right = np.concatenate([best[1:], [-np.inf]])
    idx = np.nonzero((best >= left) & (best >= right))[0]
    idx = idx[np.argsort(-best[idx], kind="stable")][:ncand]
    return idx


def run_refine(wl, plan, scores, widths, d_e, d_v, injections):
    """Refines the strongest grid peaks of a Fourier-domain FFA by gradient ascent of the
    fixed-width S/N (each candidate keeps its grid-best width); timed with events on the
    current stream.  Returns the bench's "refine" object."""
    import torch

    from loki_amd import refine

    rcfg = wl["refine"]
    nw = len(widths)
    nsamps, tsamp = wl["nsamps"], wl["tsamp"]
    t_ref = 0.5 * nsamps * tsamp
    sc = scores.view(-1, nw).cpu().numpy()
    best_w = sc.argmax(axis=1)
    best = sc.max(axis=1)
    freqs = np.asarray(plan.params_dict["freq"], np.float64)
    nf = freqs.size
    best_f = best.reshape(-1, nf).max(axis=0)      # collapse the (sub-cell) fdot/fddot axes
    arg_f = best.reshape(-1, nf).argmax(axis=0)
    idx = grid_peaks(best_f, rcfg["ncand"])
    widx = best_w.reshape(-1, nf)[arg_f[idx], idx]
    theta0 = np.zeros((idx.size, 3))
    theta0[:, 0] = freqs[idx]
    th_out = np.zeros_like(theta0)
    s0_out = np.zeros(idx.size)
    s1_out = np.zeros(idx.size)
    ev = []
    total_ms = 0.0
    evals = 0
    nbins = wl["nbins"]
    for wi in np.unique(widx):
        sel = np.nonzero(widx == wi)[0]
        dth = torch.from_numpy(theta0[sel]).to(d_e.device)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        res = refine.refine_candidates(d_e, d_v, dth, tsamp, t_ref, nbins, int(widths[wi]),
                                       iters=rcfg["iters"])
        e1.record()
        torch.cuda.synchronize()
        total_ms += e0.elapsed_time(e1)
        evals += res.evaluations * sel.size
        th_out[sel] = res.theta.cpu().numpy()
        s0_out[sel] = res.scores0.cpu().numpy()
        s1_out[sel] = res.scores.cpu().numpy()
        ev.append((int(widths[wi]), int(sel.size)))
    # one forward + one adjoint over all candidates at one width, timed on their own
    ws = torch.empty(refine.workspace_bytes(nsamps, idx.size, nbins), dtype=torch.uint8,
                     device=d_e.device)
    dth = torch.from_numpy(th_out).to(d_e.device)
    w0 = int(widths[np.bincount(widx).argmax()])
    refine.score_raw(d_e, d_v, dth, tsamp, t_ref, nbins, w0, True, ws)
    ea, eb, ec = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    ea.record()
    _, st = refine.score_raw(d_e, d_v, dth, tsamp, t_ref, nbins, w0, True, ws)
    eb.record()
    refine.grad_raw(d_e, d_v, dth, st, tsamp, t_ref, nbins, ws)
    ec.record()
    torch.cuda.synchronize()
    fwd_ms, adj_ms = ea.elapsed_time(eb), eb.elapsed_time(ec)
    terms = float(nsamps) * idx.size * (nbins // 2 + 1)
    # injections matched to the refined candidate nearest in f (within one grid cell)
    df = float(freqs[1] - freqs[0]) if nf > 1 else 1.0
    rec = []
    for sg in injections:
        j = int(np.argmin(np.abs(th_out[:, 0] - sg["f0"])))
        ok = abs(th_out[j, 0] - sg["f0"]) <= df
        rec.append((ok, s0_out[j], s1_out[j], th_out[j, 0] - sg["f0"]))
    nrec = sum(r[0] for r in rec)
    gain = [r[2] - r[1] for r in rec if r[0]]
    flops = REFINE_FLOPS_PER_TERM * terms
    return {
        "ncand": int(idx.size), "iters": rcfg["iters"], "width_groups": ev,
        "ms_total": total_ms, "evaluations": evals,
        "forward_ms": fwd_ms, "adjoint_ms": adj_ms,
        "harmonic_terms_per_pass": terms,
        "roofline": {"bound": "valu", "kernel": "refine_forward", "unit": "TFLOP/s",
                     "achieved": flops / (fwd_ms * 1e-3) / 1e12, "peak": VALU_F32_PEAK_TFLOPS,
                     "frac": flops / (fwd_ms * 1e-3) / 1e12 / VALU_F32_PEAK_TFLOPS,
                     "adjoint_achieved": flops / (adj_ms * 1e-3) / 1e12,
                     "flops_per_term": REFINE_FLOPS_PER_TERM},
        "injections_recovered": int(nrec), "injections": len(injections),
        "median_snr_gain": float(np.median(gain)) if gain else None,
        "median_abs_df_hz": float(np.median([abs(r[3]) for r in rec if r[0]])) if nrec else None,
        "snr_never_decreased": bool(np.all(s1_out >= s0_out)),
    }


def refine_cpu_baseline(wl, ts_e, ts_v) -> dict:
    """FP64 numpy adjoint (oracle/refine.py) on a bounded sample: 2 candidates over the first
    2^18 samples, in harmonic terms per second (one forward + one adjoint each)."""
    from oracle import refine as oref

    n = 1 << 18
    nbins = wl["nbins"]
    t_ref = 0.5 * n * wl["tsamp"]
    t0 = time.perf_counter()
    for f in (100.0, 300.0):
        oref.grad(ts_e[:n], ts_v[:n], wl["tsamp"], t_ref, (f, 0.0, 0.0), nbins, 4)
    dt = time.perf_counter() - t0
    terms = 2.0 * n * (nbins // 2 + 1)
    return {"value": terms / dt, "unit": "harmonic terms/s (forward + adjoint)", "cores": 1,
            "kind": "port", "sample": f"2 candidates x 2^18 samples x {nbins // 2 + 1} harmonics "
                                      f"in {dt:.2f} s (FP64 numpy)"}


def storage_error_vs_fp64(wl, ts_e, ts_v, widths) -> dict:
    """|dS/N| of every fold storage against the FP64 oracle fold on the workload's storage
    sample (same series and grid rules, a frequency sub-band around the injection)."""
    from loki_amd import libculoki, libloki
    from oracle import oracle as orc
    from oracle.plan import SearchConfig

    smp = wl["storage_sample"]
    nbins = wl["nbins"]
    ocfg = SearchConfig(wl["nsamps"], wl["tsamp"], nbins, wl["eta"], smp["param_limits"])
    orc.lib()
    t0 = time.perf_counter()
    fold64, _ = orc.compute_ffa(ts_e, ts_v, ocfg, f64=True, nthreads=cpu_share()[0])
    snr64 = orc.snr_boxcar_3d_f64(fold64, widths, nbins).reshape(-1)
    del fold64
    cpu_s = time.perf_counter() - t0
    cfg = libloki.configs.PulsarSearchConfig(
        nsamps=wl["nsamps"], tsamp=wl["tsamp"], nbins=nbins, eta=wl["eta"],
        param_limits=np.array(smp["param_limits"]), use_fourier=False)
    ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
    top = np.argsort(snr64)[-100:]
    out = {"sample": f"param_limits {smp['param_limits']}: {snr64.size // len(widths)} trials "
                     f"x {len(widths)} widths (FP64 oracle fold in {cpu_s:.1f} s)",
           "max_fp64_snr": float(snr64.max())}
    for st in ("f32", "f16", "bf16", "bfp16"):
        ffa.set_fold_storage(st)
        got = ffa.scores(ts_e, ts_v, widths).astype(np.float64)
        d = np.abs(got - snr64)
        out[st] = {"max": float(d.max()), "p99": float(np.percentile(d, 99)),
                   "median": float(np.median(d)), "top100_max": float(d[top].max()),
                   "frac_below_1e-3": float(np.mean(d < 1e-3))}
    return out


def run_sweep(args, wl) -> int:
    """The sweep workload: value = final trials of all chunks per second, end to end (host
    series in, candidates and the result archive out)."""
    import tempfile

    from loki_amd import cands, libculoki, libloki
    from loki_amd.synth import make_injected_series

    cfg = libloki.configs.PulsarSearchConfig(
        nsamps=wl["nsamps"], tsamp=wl["tsamp"], nbins=wl["nbins"], eta=wl["eta"],
        param_limits=np.array(wl["param_limits"]), use_fourier=False, snr_min=wl["snr_min"],
        max_process_memory_gb=wl["max_process_memory_gb"])
    ts_e, ts_v, _ = make_injected_series(wl["nsamps"], wl["tsamp"], wl["seed"],
                                         wl["injections"], nbins=wl["nbins"])
    sweep = libculoki.ffa.FFAFreqSweepCUDA(cfg, device_id=0)
    planner = sweep.planner
    trials = sum(int(libloki.plans.FFAPlanTime(c).ncoords[-1]) for c in planner.cfgs)
    with tempfile.TemporaryDirectory() as out:
        for _ in range(args.warmup):
            sweep.execute(ts_e, ts_v, out, "bench")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            path = sweep.execute(ts_e, ts_v, out, "bench")
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        res = cands.read_results(path)
    snr, rows = res["snr"], res["param_sets"]
    found = []
    for sg in wl["injections"]:
        near = np.abs(rows[:, -2] - sg["f0"]) < 0.05
        found.append(float(snr[near].max()) if near.any() else None)
    result = {
        "metric": "FFA param trials/s (frequency sweep)", "value": trials / (ms * 1e-3),
        "unit": "param trials/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "nsamps": wl["nsamps"], "tsamp": wl["tsamp"],
                   "param_limits": wl["param_limits"], "snr_min": wl["snr_min"],
                   "max_process_memory_gb": wl["max_process_memory_gb"],
                   "chunks": [{"nbins": int(c.nbins), "eta": float(c.eta),
                               "freq": [float(x) for x in c.param_limits[-1]]}
                              for c in planner.cfgs],
                   "trials_per_step": trials},
        "roofline": None,
        "candidates": int(snr.size),
        "per_chunk_candidates": [int(x) for x in sweep.per_region],
        "injected_snr_found": found,
    }
    print(json.dumps(result), flush=True)
    return 0


def run_batch(args, wl) -> int:
    """The batched workload: value = final trials of all series per second."""
    import torch

    from loki_amd import libculoki, libloki
    from loki_amd.synth import make_batch_series, make_batch_specs

    specs = make_batch_specs(wl["nseries"], wl["seed"])
    cfgs = [libloki.configs.PulsarSearchConfig(
        nsamps=sp["nsamps"], tsamp=wl["tsamp"], nbins=sp["nbins"], eta=wl["eta"],
        param_limits=np.array([[sp["f_lo"], sp["f_hi"]]]), use_fourier=False) for sp in specs]
    widths = np.array(WIDTHS, np.uint64)
    dev = torch.device("cuda:0")
    batch = libculoki.ffa.FFABatchCUDA(cfgs, widths=widths)
    for i, sp in enumerate(specs):
        e, v = make_batch_series(sp, wl["tsamp"])
        batch.set_input(i, e, v)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.synchronize()
    schedule = batch.autotune(repeats=2, stream=stream.cuda_stream)
    info = batch.info()
    # graphs off: the same launches enqueued one by one (reported beside the graph rate)
    batch.set_graph(False)
    with torch.cuda.stream(stream):
        batch.execute(stream.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.execute(stream.cuda_stream)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    batch.set_graph(True)
    for _ in range(max(args.warmup, 1)):
        batch.execute(stream.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.execute(stream.cuda_stream)
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    trials = int(info["total_trials"])
    # algorithmic bytes of the whole batch (per-series pipeline bytes, FP32 time folds), and
    # the level-0 brute fold's (sample, frequency) pairs: each reads one staged (e, v)
    # float pair from LDS
    alg = 0
    pairs = 0
    for i in range(batch.size):
        plan = batch.plan(i)
        nb = int(cfgs[i].nbins)
        alg += pipeline_bytes(plan, nb, int((widths < nb).sum()), int(cfgs[i].nsamps))
        pairs += int(cfgs[i].nsamps) * int(plan.ncoords[0])
    lds_bytes = 8 * pairs
    # recovery: best S/N of the injected series at its true period
    hits = 0
    for i, sp in enumerate(specs):
        if sp["signal"] is None:
            continue
        nw = int((widths < sp["nbins"]).sum())
        sc = batch.scores(i).reshape(-1, nw)
        grid = batch.plan(i).params_dict["freq"]
        best = int(np.argmax(sc.max(axis=1)))
        df = grid[1] - grid[0] if len(grid) > 1 else 0.0
        hits += int(abs(grid[best] - sp["signal"]["f0"]) <= 2 * abs(df) + 1e-12)
    nsig = sum(1 for sp in specs if sp["signal"] is not None)
    result = {
        "metric": "FFA param trials/s (batched uneven series)",
        "value": trials / (ms_per_step * 1e-3),
        "unit": "param trials/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload, "nseries": len(specs),
            "nsamps_range": [min(sp["nsamps"] for sp in specs), max(sp["nsamps"] for sp in specs)],
            "nbins": sorted({sp["nbins"] for sp in specs}), "trials_per_step": trials,
            "launches_per_step": int(info["launches"]), "device_bytes": int(info["device_bytes"]),
            "schedule": schedule,
        },
        "roofline": {
            "bound": "lds", "kernel": "k_bf_units_grouped (level-0 brute fold)",
            "achieved": lds_bytes / (ms_per_step * 1e-3) / 1e12, "peak": LDS_PEAK_TBPS,
            "unit": "TB/s (LDS)",
            "frac": lds_bytes / (ms_per_step * 1e-3) / 1e12 / LDS_PEAK_TBPS, "traffic": None,
            "lds_bytes_per_step": lds_bytes, "sample_freq_pairs": pairs,
            "note": "8 LDS bytes per (sample, frequency) pair of the brute fold, which takes 95 % "
                    "of the step (profiles/batched_uneven_kernel_trace.md; its SQ counters in "
                    "profiles/batched_uneven_brute_fold_sq_counters.md), over the whole step's "
                    "time, so the fold's own fraction is about 5 % higher; LDS peak 128 B per "
                    "clock per CU",
            "hbm": {"alg_bytes_per_step": alg,
                    "achieved_GBps": alg / (ms_per_step * 1e-3) / 1e9,
                    "frac": alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS},
        },
        "graph_vs_eager_ms": {"graph": ms_per_step, "eager": eager_ms},
        "recovered": {"signals": nsig, "found_within_2_cells": hits},
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        from oracle.plan import Plan, SearchConfig

        threads = args.cpu_threads or cpu_share()[0]
        sub = specs[: wl["cpu_sample"]["nseries"]]
        ntr, dt = 0, 0.0
        orc.lib()
        for sp in sub:
            ocfg = SearchConfig(sp["nsamps"], wl["tsamp"], sp["nbins"], wl["eta"],
                                [(sp["f_lo"], sp["f_hi"])], use_fourier=False)
            e, v = make_batch_series(sp, wl["tsamp"])
            w = widths[widths < sp["nbins"]]
            ta = time.perf_counter()
            orc.compute_ffa_scores(e, v, ocfg, widths=w, nthreads=threads)
            dt += time.perf_counter() - ta
            ntr += Plan(ocfg).ncoords[-1]
        result["cpu_baseline"] = {
            "value": ntr / dt, "unit": "param trials/s", "cores": threads, "kind": "port",
            "sample": f"the first {len(sub)} series of the batch, one after another: {ntr} final "
                      f"trials in {dt:.2f} s",
            **cpu_share()[1],
        }
    print(json.dumps(result), flush=True)
    return 0


def run_seeds(dev, stream, steps: int) -> dict:
    """Pruning seeds (kernels_cuda.cu:16-277) over synthetic final-level rows in HBM: 2^21
    leaves, nbins 64, a final fold of 8 segments x 2^16 coordinates.  Per launch, the linear
    seed moves 24 nbins + 13 bytes per leaf (tree row, rotated FFA row, output row, two indices,
    shift, mask) and the ascending seed 8 nbins (G + 1) + 8 G bytes for G segments; rows that
    several leaves share may come from the caches, so the minimum HBM bytes (every distinct
    row once) are given beside them."""
    import ctypes

    import torch

    import loki_amd

    lib = ctypes.CDLL(str(loki_amd.LIBRARY_PATH))
    vp, sz, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32
    lib.loki_shift_add_linear_batch_device.argtypes = [vp] * 7 + [sz, sz, u32, u32, sz, vp]
    lib.loki_shift_add_ascend_linear_batch_device.argtypes = [vp] * 5 + [sz] * 5 + [vp]
    lib.loki_shift_add_linear_complex_batch_device.argtypes = [vp] * 7 + [sz, sz, sz, u32, u32,
                                                                         sz, vp]
    nbins, nleaves, G, nci = 64, 1 << 21, 8, 1 << 16
    nbf = nbins // 2 + 1
    g = torch.Generator(device=dev).manual_seed(7)
    tree = torch.randn(nleaves, 2, nbins, device=dev, generator=g)
    ffa = torch.randn(G, nci, 2, nbins, device=dev, generator=g)
    out = torch.empty(nleaves, 2, nbins, device=dev)
    # a permutation: every tree row is read once
    itree = torch.randperm(nleaves, device=dev, generator=g).to(torch.int32)
    iffa = torch.randint(0, G * nci, (nleaves,), device=dev, dtype=torch.int32, generator=g)
    iffa_a = torch.randint(0, nci, (G, nleaves), device=dev, dtype=torch.int32, generator=g)
    iseg = torch.arange(G, device=dev, dtype=torch.int32)
    shift = torch.rand(nleaves, device=dev, generator=g) * nbins
    shift_a = torch.rand(G, nleaves, device=dev, generator=g) * nbins
    mask = torch.ones(nleaves, device=dev, dtype=torch.uint8)
    tree_c = torch.randn(nleaves, 2, nbf, 2, device=dev, generator=g)
    ffa_c = torch.randn(G * nci, 2, nbf, 2, device=dev, generator=g)
    out_c = torch.empty(nleaves, 2, nbf, 2, device=dev)
    torch.cuda.synchronize()
    s = stream.cuda_stream

    def linear():
        return lib.loki_shift_add_linear_batch_device(
            tree.data_ptr(), itree.data_ptr(), mask.data_ptr(), ffa.data_ptr(), iffa.data_ptr(),
            shift.data_ptr(), out.data_ptr(), nbins, nleaves, 0, nleaves, G * nci, s)

    def ascend():
        return lib.loki_shift_add_ascend_linear_batch_device(
            ffa.data_ptr(), iseg.data_ptr(), iffa_a.data_ptr(), shift_a.data_ptr(),
            out.data_ptr(), nbins, nci, nleaves, G, G, s)

    def linear_c():
        return lib.loki_shift_add_linear_complex_batch_device(
            tree_c.data_ptr(), itree.data_ptr(), mask.data_ptr(), ffa_c.data_ptr(),
            iffa.data_ptr(), shift.data_ptr(), out_c.data_ptr(), nbf, nbins, nleaves, 0,
            nleaves, G * nci, s)

    res = {"nbins": nbins, "n_leaves": nleaves, "segments": G, "ffa_rows": G * nci}
    ffa_rows_b = G * nci * 2 * nbins * 4
    for name, fn, logical, hbm_min in (
        ("linear", linear, nleaves * (24 * nbins + 13),
         nleaves * (16 * nbins + 13) + min(ffa_rows_b, nleaves * 8 * nbins)),
        ("ascend", ascend, nleaves * (8 * nbins * (G + 1) + 8 * G),
         nleaves * (8 * nbins + 8 * G) + ffa_rows_b),
        ("linear_complex", linear_c, nleaves * (48 * nbf + 13),
         nleaves * (32 * nbf + 13) + min(G * nci * 16 * nbf, nleaves * 16 * nbf)),
    ):
        assert fn() == 0, lib.loki_last_error()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / steps
        res[name] = {"avg_ms": round(ms, 4), "logical_bytes": logical,
                     "logical_GBps": round(logical / (ms * 1e-3) / 1e9, 1),
                     "hbm_min_bytes": hbm_min,
                     "hbm_min_GBps": round(hbm_min / (ms * 1e-3) / 1e9, 1),
                     "hbm_min_frac": round(hbm_min / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 3)}
    del tree, ffa, out, tree_c, ffa_c, out_c
    torch.cuda.empty_cache()
    return res


def stored_level_bytes(plan, nbins, level, storage):
    """Bytes of fold level `level` as stored (time domain).  Block floating point: int16
    mantissas plus one 16-bit exponent word per row pair, and level 0 in FP32."""
    n = int(np.prod([int(x) for x in plan.fold_shapes[level]]))
    if storage == "bfp16":
        return 4 * n if level == 0 else 2 * n + n // nbins
    return n * (4 if storage == "f32" else 2)


def fused_roofline(plan, nbins, nw, args, wl, avg_ms, dom_name, variant, traffic):
    """Roofline of the fused last level.  It is VALU-issue bound: `frac` is its VALU issue
    rate, SQ_INSTS_VALU per launch (rocprofv3 --pmc, profiles/valu_counters.json) x 64 lanes
    over this run's average launch time, against the dual-issue VALU peak (1024 SIMDs x 2
    wave64 instructions per 4 cycles x 2.4 GHz = the 157.3 TFLOP/s FP32 vector peak counted
    as FMAs); `single_issue_frac` is the same rate against one instruction per 4 cycles.  Its
    HBM side is reported beside it: the minimum bytes (the stored level it reads, once, plus
    the scores) and the measured traffic, each over the same time."""
    nl = int(plan.n_levels)
    depth = args.fuse_depth if nbins in (8, 16, 32) and nl >= 3 else 1
    src = nl - 1 - depth
    nc = int(plan.ncoords[-1])
    min_bytes = stored_level_bytes(plan, nbins, src, args.storage) + nc * nw * 4
    t = avg_ms * 1e-3
    hbm = {"min_bytes": min_bytes, "min_GBps": min_bytes / t / 1e9,
           "min_frac": min_bytes / t / 1e9 / HBM_PEAK_GBPS,
           "reads_level": src, "traffic": traffic,
           "traffic_frac": (traffic / t / 1e9 / HBM_PEAK_GBPS) if traffic else None,
           "traffic_source": "profiles/hbm_traffic.json (FETCH_SIZE x 2 + WRITE_SIZE)"
                             if traffic else None}
    ctr = load_counters(args.workload, variant)
    if ctr and ctr.get("SQ_INSTS_VALU"):
        lane_inst = float(ctr["SQ_INSTS_VALU"]) * 64
        achieved = lane_inst / t / 1e12
        return {"bound": "valu", "kernel": dom_name, "achieved": achieved,
                "peak": VALU_LANE_INST_PEAK_T, "unit": "T lane-instructions/s",
                "frac": achieved / VALU_LANE_INST_PEAK_T,
                "single_issue_frac": 2.0 * achieved / VALU_LANE_INST_PEAK_T, "traffic": traffic,
                "avg_ms": avg_ms, "valu_lane_inst_per_launch": lane_inst,
                "valu_lane_inst_per_trial": lane_inst / nc,
                "counters": {k: v for k, v in ctr.items()},
                "counters_source": "profiles/valu_counters.json", "hbm": hbm}
    # without a counter profile for this variant: the HBM view on minimum bytes
    return {"bound": "hbm", "kernel": dom_name, "achieved": hbm["min_GBps"],
            "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": hbm["min_frac"], "traffic": traffic,
            "avg_ms": avg_ms, "alg_bytes_per_launch": min_bytes,
            "note": "fused last level: minimum bytes (stored level read once + scores)",
            "hbm": hbm}


def practical_hbm(nbytes: int = 2 << 30) -> dict:
    """What a plain streaming kernel reaches on this GPU: torch's device-to-device copy of
    nbytes (read + write), best of 3 by hipEvent, beside the 8 TB/s peak."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    out = {}
    for name, fn, moved in (("copy", lambda: b.copy_(a), 2 * nbytes),):
        fn()
        best = None
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        out[f"{name}_GBps"] = moved / (best * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    out["peak_GBps"] = HBM_PEAK_GBPS
    out["note"] = ("a plain streaming copy (torch copy_ of 2 GiB, bytes read + written) in the "
                   "same run: the bandwidth a merge level can expect in practice, beside the "
                   "nominal peak")
    return out


def step_floor(plan, nbins, nw, nsamps, args, ms_per_step, fourier):
    """HBM floor of one step: the input series read once, every level that goes through HBM
    written once and read once, the scores written -- against the measured step."""
    if fourier:
        return None
    nl = int(plan.n_levels)
    depth = args.fuse_depth if nbins in (8, 16, 32) and nl >= 3 else 1
    if args.unfused:
        depth = 0
    top = nl - 1 - depth
    b = 2 * nsamps * 4 + int(plan.ncoords[-1]) * nw * 4
    for lvl in range(0, top + 1):
        st = args.storage if lvl < nl - 1 else "f32"
        b += 2 * stored_level_bytes(plan, nbins, lvl, st)
    floor_ms = b / (HBM_PEAK_GBPS * 1e9) * 1e3
    return {"min_bytes": b, "floor_ms": floor_ms, "frac": floor_ms / ms_per_step,
            "levels_through_hbm": top + 1}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs the FFA hot path on one GPU; --gpus {args.gpus} is "
              "not supported (multi-GPU sharding is design-only).", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    import torch

    from loki_amd import libculoki, libloki
    from loki_amd.synth import make_series

    wl = WORKLOADS[args.workload]
    if wl.get("batch"):
        return run_batch(args, wl)
    if wl.get("sweep"):
        return run_sweep(args, wl)
    nbins = wl["nbins"]
    widths = np.array([w for w in WIDTHS if w < nbins], dtype=np.uint64)
    sig = wl["signal"]
    injections = None
    if "injections" in wl:
        from loki_amd.synth import make_injected_series, make_injections

        inj = wl["injections"]
        injections = make_injections(inj["n"], inj["seed"], inj["f"], inj["fdot"], inj["fddot"],
                                     inj["snr"], inj["fwhm"])
        ts_e, ts_v, _amps = make_injected_series(wl["nsamps"], wl["tsamp"], wl["seed"],
                                                 injections, nbins=nbins, widths=widths,
                                                 device=torch.device("cuda:0"))
    else:
        ts_e, ts_v, _amp = make_series(wl["nsamps"], wl["tsamp"], wl["seed"], f0=sig["f0"],
                                       accel=sig["accel"], jerk=sig["jerk"], fwhm=sig["fwhm"],
                                       snr=sig["snr"], nbins=nbins, widths=widths)
    fourier = bool(wl.get("fourier", False))
    cfg = libloki.configs.PulsarSearchConfig(
        nsamps=wl["nsamps"], tsamp=wl["tsamp"], nbins=nbins, eta=wl["eta"],
        param_limits=np.array(wl["param_limits"]), use_fourier=fourier)
    dev = torch.device("cuda:0")
    if fourier:
        ffa = libculoki.ffa.FFACOMPLEXCUDA(cfg, device_id=0)
        ffa.set_fused(not args.unfused)
        ffa.set_dft_precision(args.dft)
    else:
        ffa = libculoki.ffa.FFACUDA(cfg, device_id=0)
        ffa.set_exact_norm(not args.fast_norm)
        ffa.set_fused(not args.unfused)
        ffa.set_fuse_depth(args.fuse_depth)
        ffa.set_fold_storage(args.storage)
    plan = ffa.plan
    nc = int(plan.ncoords[-1])
    nw = int(widths.size)
    d_e = torch.from_numpy(ts_e).to(dev)
    d_v = torch.from_numpy(ts_v).to(dev)
    scores = torch.empty(nc * nw, dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(device=dev)

    def step():
        ffa.scores_device(d_e.data_ptr(), d_v.data_ptr(), widths, scores.data_ptr(), nc * nw,
                          False, stream.cuda_stream)

    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            step()
    torch.cuda.synchronize()

    # timed region: exactly `steps` steps, no per-kernel events
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for _ in range(args.steps):
            step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps

    # per-kernel durations: hipEvent pairs on the launch stream, separate steps
    ffa.set_timing(True)
    per_kernel = defaultdict(list)
    with torch.cuda.stream(stream):
        for _ in range(min(args.steps, 10)):
            step()
            for kind, level, ms in ffa.timings():  # waits for this step's events
                per_kernel[(kind, level)].append(ms)
    torch.cuda.synchronize()
    ffa.set_timing(False)

    def merge_levels_ms():
        """hipEvent time of the stored levels' merges (the memory-bound part that the fold
        storage changes), min over 3 steps."""
        ffa.set_timing(True)
        best = None
        with torch.cuda.stream(stream):
            for _ in range(3):
                step()
                t = sum(ms for kind, _level, ms in ffa.timings() if str(kind) == "merge")
                best = t if best is None else min(best, t)
        torch.cuda.synchronize()
        ffa.set_timing(False)
        return best

    # The same step with the other fold storage precisions (same series, same stream).
    storage_cmp = None
    if not fourier and not args.no_storage_compare:
        storage_cmp = {args.storage: ms_per_step}
        merge_cmp = {args.storage: merge_levels_ms()}
        for st in ("f32", "f16", "bf16", "bfp16"):
            if st == args.storage:
                continue
            ffa.set_fold_storage(st)
            with torch.cuda.stream(stream):
                step()
            torch.cuda.synchronize()
            ta = time.perf_counter()
            with torch.cuda.stream(stream):
                for _ in range(args.steps):
                    step()
            torch.cuda.synchronize()
            storage_cmp[st] = (time.perf_counter() - ta) * 1e3 / args.steps
            merge_cmp[st] = merge_levels_ms()
        ffa.set_fold_storage(args.storage)
        with torch.cuda.stream(stream):
            step()  # scores of the main run's storage for the recovery check
        torch.cuda.synchronize()
        storage_cmp = {
            "ms_per_step": {k: round(v, 4) for k, v in storage_cmp.items()},
            "speedup_vs_f32": {k: round(storage_cmp["f32"] / v, 4)
                               for k, v in storage_cmp.items()},
            "merge_levels_ms": {k: round(v, 4) for k, v in merge_cmp.items()},
            "merge_levels_speedup_vs_f32": {k: round(merge_cmp["f32"] / v, 4)
                                            for k, v in merge_cmp.items() if v > 0},
            "note": "stored levels at that precision (bfp16: block floating point, level 0 in "
                    "FP32), every add in FP32; merge_levels_ms: the stored levels' merges alone "
                    "(hipEvent, the memory-bound part), the fused last level reads them once "
                    "and computes in FP32 for every storage; S/N error vs FP64 in "
                    "snr_error_vs_fp64 (jerk workload) and "
                    "tests/test_gpu_parity.py::test_storage_snr_error_vs_fp64",
        }

    # The same step with the hardware reciprocal square root in the profile normalisation
    # (|dS/N| <= 1e-4 against the exact scores, tests/test_gpu_parity.py and
    # tests/test_gpu_configs.py), beside the bit-exact default.
    norm_cmp = None
    if not fourier and not args.fast_norm:
        ffa.set_exact_norm(False)
        with torch.cuda.stream(stream):
            step()
        torch.cuda.synchronize()
        tn0 = time.perf_counter()
        with torch.cuda.stream(stream):
            for _ in range(args.steps):
                step()
        torch.cuda.synchronize()
        n_ms = (time.perf_counter() - tn0) * 1e3 / args.steps
        ffa.set_exact_norm(True)
        with torch.cuda.stream(stream):
            step()  # exact scores again for the recovery check
        torch.cuda.synchronize()
        norm_cmp = {"exact_ms_per_step": round(ms_per_step, 4),
                    "hw_rsqrt_ms_per_step": round(n_ms, 4),
                    "hw_rsqrt_param_trials_per_s": nc / (n_ms * 1e-3),
                    "note": "exact: correctly rounded e * (1/sqrt(v)), bit-identical to the CPU "
                            "oracle; hw_rsqrt: e * rsqrt(v), |dS/N| <= 1e-4"}

    # Score-and-filter (the FFAFreqSweep step, ffa_freq_sweep_cuda.cu:219-260): the same FFA with
    # only the (trial, width) pairs above the threshold leaving the last-level kernel.
    score_filter = None
    if not args.no_score_filter:
        cap = 1 << 24
        f_idx = torch.empty(cap, dtype=torch.int32, device=dev)
        f_snr = torch.empty(cap, dtype=torch.float32, device=dev)
        f_cnt = torch.zeros(1, dtype=torch.int32, device=dev)

        def fstep():
            ffa.scores_filter_device(d_e.data_ptr(), d_v.data_ptr(), widths,
                                     args.filter_threshold, False, f_idx.data_ptr(),
                                     f_snr.data_ptr(), cap, f_cnt.data_ptr(), stream.cuda_stream)

        with torch.cuda.stream(stream):
            fstep()
        torch.cuda.synchronize()
        tf0 = time.perf_counter()
        with torch.cuda.stream(stream):
            for _ in range(args.steps):
                fstep()
        torch.cuda.synchronize()
        f_ms = (time.perf_counter() - tf0) * 1e3 / args.steps
        score_filter = {"threshold": args.filter_threshold, "ms_per_step": f_ms,
                        "param_trials_per_s": nc / (f_ms * 1e-3),
                        "candidates": int(f_cnt.item()), "capacity": cap}

    seeds = None
    if not args.no_seeds and not fourier:
        seeds = run_seeds(dev, stream, args.steps)

    # signal recovery check (outside the timed region)
    best = int(torch.argmax(scores).item())
    shape = [int(x) for x in plan.fold_shapes[-1][1:-2]]
    idx = np.unravel_index(best // nw, shape)
    grid = plan.params_dict
    found = {name: float(grid[name][i]) for name, i in zip(cfg.param_names, idx)}
    found["snr"] = float(scores[best].item())
    found["width"] = int(widths[best % nw])

    # dominant kernel and its roofline
    avg = {k: float(np.mean(v)) for k, v in per_kernel.items()}
    dom = max(avg, key=avg.get)
    dom_name = f"{dom[0]}_L{dom[1]}"
    variant = dom_name if dom[0] != "merge_score" else (
        f"{dom_name}_depth{args.fuse_depth}_{args.storage}"
        f"{'' if not args.fast_norm else '_hwrsqrt'}")
    traffic = load_traffic(args.workload, variant)
    if fourier and dom[0] == "brute_fold":
        # the DFT brute fold runs on the matrix cores: FP32 MFMA roofline
        work = dft_flops(plan, nbins, wl["nsamps"])
        roof = {"bound": "mfma", "kernel": dom_name,
                "achieved": work / (avg[dom] * 1e-3) / 1e12, "peak": MFMA_F32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "traffic": traffic, "flops_per_launch": work,
                "avg_ms": avg[dom]}
        roof["frac"] = roof["achieved"] / MFMA_F32_PEAK_TFLOPS
    elif dom[0] == "merge_score":
        roof = fused_roofline(plan, nbins, nw, args, wl, avg[dom], dom_name, variant, traffic)
    else:
        dom_bytes = alg_bytes(plan, nbins, nw, dom[0], dom[1], wl["nsamps"], fourier)
        achieved = dom_bytes / (avg[dom] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom_name, "achieved": achieved,
                "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                "traffic": traffic, "alg_bytes_per_launch": dom_bytes, "avg_ms": avg[dom]}
    kern_ms = {f"{k}_L{l}": round(v, 4) for (k, l), v in sorted(avg.items(), key=lambda x: -x[1])}
    total_alg = pipeline_bytes(plan, nbins, nw, wl["nsamps"], fourier)

    result = {
        "metric": METRIC,
        "value": nc / (ms_per_step * 1e-3),
        "unit": "param trials/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64" if fourier else "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload, "nsamps": wl["nsamps"], "tsamp": wl["tsamp"],
            "nbins": nbins, "eta": wl["eta"], "param_limits": wl["param_limits"],
            "widths": [int(w) for w in widths], "fold_final": [int(x) for x in plan.fold_shapes[-1]],
            "n_levels": int(plan.n_levels), "trials_per_step": nc,
            "domain": "fourier" if fourier else "time",
            "fold_storage": "c64" if fourier else args.storage,
            "dft": args.dft if fourier else None,
            "score_norm": "hw_rsqrt" if args.fast_norm else "exact",
            "final_level": ("unfused" if args.unfused else
                            "fused_irfft_score" if fourier else "fused"),
        },
        "roofline": roof,
        "hbm_practical": practical_hbm() if not fourier else None,
        "step_floor": step_floor(plan, nbins, nw, wl["nsamps"], args, ms_per_step, fourier),
        "unfused_alg_bytes": total_alg,
        "kernel_ms": kern_ms,
        "storage_compare": storage_cmp,
        "norm_compare": norm_cmp,
        "score_filter": score_filter,
        "pruning_seeds": seeds,
        "found": found,
        "injected": {"freq": sig["f0"], "accel": sig["accel"], "snr": sig["snr"]},
    }
    if storage_cmp is not None and "storage_sample" in wl:
        storage_cmp["snr_error_vs_fp64"] = storage_error_vs_fp64(wl, ts_e, ts_v, widths)
    if injections is not None:
        result["injected"] = {"n": len(injections), "spec": wl["injections"]}
        if not args.no_refine:
            result["refine"] = run_refine(wl, plan, scores, widths, d_e, d_v, injections)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or cpu_share()[0]
        result["cpu_baseline"] = cpu_baseline(wl, ts_e, ts_v, widths, threads)
        if "refine" in result:
            result["refine"]["cpu_baseline"] = refine_cpu_baseline(wl, ts_e, ts_v)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
