"""CPU oracle for the fast-folding (FFA) hot path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference's CPU
algorithm (pravirkr/loki v0.0.3) so that the HIP product path can be checked
against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import, call, link or execute
anything in here, and only as the checker / the timed CPU baseline -- never as
the product path.  The product (``loki_amd``) never imports this package and
fails loudly when its HIP library is missing.

Contents
--------
plan.py      -- config / plan geometry, grid helpers and the FP64 shift-table
                resolvers (numpy, vectorised), following configs.cpp,
                psr_utils.{hpp,cpp}, plans.cpp and taylor_ffa.cpp.
liboracle.c  -- C (OpenMP) restatement of the kernels: bucket brute fold,
                time-domain merges, FP64 Fourier brute fold / merges, textbook
                inverse real DFT and boxcar S/N (kernels.cpp, fold.cpp,
                score.cpp, utils.cpp).  Built by ``oracle/Makefile`` into
                ``oracle/liboracle.so``.
oracle.py    -- ctypes wrappers + whole-pipeline drivers mirroring ffa.cpp.

Every function's docstring / comment cites the reference file:line it follows
and the test that pins it (tests/test_oracle_*.py).
"""
