"""Write the small seeded fixtures in tests/golden/ffa_small.npz.

The reference holds no golden vectors for the fold / FFA / score path and cannot be built or
imported here, so these fixtures come from the oracle (oracle/).  The oracle is pinned
separately, in tests/test_oracle.py.  The fixtures freeze its outputs on small inputs so that:
  * a later change to the oracle that alters any bit fails tests/test_golden.py;
  * GPU runs compare against committed data, independently of the oracle build.

Run from the repository root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

TSAMP = 64e-6
# name: (nsamps, nbins, param_limits, widths)
CASES = {
    "freq": (2**14, 16, [(8.0, 12.0)], [1, 2, 3, 4, 6]),
    "accel": (2**15, 16, [(-1.0e6, 1.0e6), (60.0, 64.0)], [1, 2, 4, 8]),
    "jerk": (2**15, 16, [(-3.0e7, 3.0e7), (-1.0e6, 1.0e6), (60.0, 62.0)], [1, 3, 5]),
}


def make() -> dict:
    from oracle import oracle as orc
    from oracle.plan import SearchConfig

    out = {}
    for i, (name, (nsamps, nbins, lims, widths)) in enumerate(CASES.items()):
        rng = np.random.default_rng(4242 + i)
        ts_e = rng.standard_normal(nsamps, dtype=np.float32)
        ts_v = (1.0 + 0.25 * rng.random(nsamps)).astype(np.float32)
        w = np.asarray(widths, np.uint64)
        cfg_t = SearchConfig(nsamps, TSAMP, nbins, 1.0, lims, use_fourier=False)
        fold_t, _ = orc.compute_ffa(ts_e, ts_v, cfg_t, nthreads=1)
        scores_t, _ = orc.compute_ffa_scores(ts_e, ts_v, cfg_t, widths=w, nthreads=1)
        cfg_f = SearchConfig(nsamps, TSAMP, nbins, 1.0, lims, use_fourier=True)
        fold_f, _ = orc.compute_ffa(ts_e, ts_v, cfg_f, fourier=True, nthreads=1)
        out[f"{name}_ts_e"] = ts_e
        out[f"{name}_ts_v"] = ts_v
        out[f"{name}_param_limits"] = np.asarray(lims, np.float64)
        out[f"{name}_nbins"] = np.int64(nbins)
        out[f"{name}_widths"] = w
        out[f"{name}_fold_time"] = fold_t
        out[f"{name}_scores_time"] = scores_t
        out[f"{name}_fold_fourier"] = fold_f
    return out


def main() -> None:
    from oracle import oracle as orc

    orc.build()
    data = make()
    path = Path(__file__).resolve().parent / "ffa_small.npz"
    np.savez_compressed(path, **data)
    print(f"wrote {path} ({path.stat().st_size} bytes, {len(data)} arrays)")


if __name__ == "__main__":
    main()
