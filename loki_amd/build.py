"""Build the loki_amd native libraries in-tree (no JIT cache, no pip install).

Outputs (git-ignored, shipped with the tree):
  loki_amd/libloki_amd.so                        -- C-ABI + gfx950 HIP kernels
  loki_amd/_loki_amd.cpython-*.so                -- pybind11 bindings over the C-ABI

Usage:  python loki_amd/build.py   (or __graft_entry__.build()); run it as a script, since
importing the loki_amd package needs the libraries this builds.
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = ROOT / "_build" / "obj"
ARCH = os.environ.get("LOKI_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["k_resolve.hip", "k_brutefold.hip", "k_merge.hip", "k_score.hip", "k_final.hip", "k_final2.hip",
               "k_fourier.hip", "k_fourier_score.hip", "k_refine.hip", "k_seed.hip", "k_tree.hip",
               "ffa.hip", "batch.hip", "sweep.hip", "capi.hip"]
HOST_SOURCES = ["config.cpp", "plan.cpp", "regions.cpp"]

# -ffp-contract=off: no FMA contraction anywhere (bit-exact FP64 shift tables
# and FP32 score arithmetic).  Packed FP32 instructions are disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
]
HOST_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    # surface warnings other than the known host-pass feature notice
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l
                    and "is not a recognized feature" not in l and l.strip())
    if err:
        sys.stderr.write(err + "\n")


def _compile_hip(src: str) -> Path:
    out = OBJ / (src + ".o")
    _run([HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src), "-o", str(out)])
    return out


def _compile_host(src: str) -> Path:
    out = OBJ / (src + ".o")
    _run(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(out)])
    return out


def ext_suffix() -> str:
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build(verbose: bool = True) -> None:
    OBJ.mkdir(parents=True, exist_ok=True)
    jobs = int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, 16))) as ex:
        futs = [ex.submit(_compile_hip, s) for s in HIP_SOURCES]
        futs += [ex.submit(_compile_host, s) for s in HOST_SOURCES]
        objs = [f.result() for f in futs]
    lib = PKG / "libloki_amd.so"
    tmp = PKG / "libloki_amd.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, lib)
    # pybind11 bindings (host-only C++, links the C-ABI library)
    import pybind11

    pyinc = sysconfig.get_paths()["include"]
    mod = PKG / ("_loki_amd" + ext_suffix())
    tmpm = PKG / ("_loki_amd.tmp" + ext_suffix())
    _run(["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-fvisibility=hidden",
          "-I", pybind11.get_include(), "-I", pyinc, "-I", str(ROOT / "include"),
          str(CSRC / "pymodule.cpp"), "-L", str(PKG), "-lloki_amd",
          "-Wl,-rpath,$ORIGIN", "-o", str(tmpm)])
    os.replace(tmpm, mod)
    if verbose:
        print(f"built {lib.relative_to(ROOT)} and {mod.relative_to(ROOT)}")


if __name__ == "__main__":
    build()
