"""The benchmark's accounting helpers on inputs with known answers, and the measurement
library's exports (no GPU calls)."""

import ctypes
import os
from types import SimpleNamespace

import numpy as np

import bench


def test_offdiag_fraction_counts_nearest_image_pairs_below_the_table_end():
    L = 10.0
    # three particles: (0,1) at 2.0, (0,2) at 9.0 -> nearest image 1.0, (1,2) at 7.0 -> 3.0
    p = np.zeros((5, 3, 3))
    p[:, 1, 0] = 2.0
    p[:, 2, 0] = 9.0
    assert bench.offdiag_fraction([p], (L, L, L), 2.5) == 2 / 3
    assert bench.offdiag_fraction([p], (L, L, L), 3.5) == 1.0
    assert bench.offdiag_fraction([p], (L, L, L), 0.5) == 0.0


def test_lds_bytes_per_launch():
    st = SimpleNamespace(pair_evals=1000, stages=10)
    # mixed: (24 + 16 (1 + 5 f)) per evaluation + 16 B per staged row (nint x ncol) per stage
    got = bench.lds_bytes_per_launch(st, 2, True, 0.5, 199, 6, 5)
    assert got == (1000 * (24 + 16 * 3.5) + 10 * 199 * 6 * 16) / 2
    got64 = bench.lds_bytes_per_launch(st, 1, False, 0.0, 199, 6, 5)
    assert got64 == 1000 * (24 + 32) + 10 * 199 * 6 * 32


def test_bytes_per_launch_matches_the_stage_formula():
    # one stage of K = 2 links for N = 4: 2 (N - 1) K = 12 evaluations, 24 (K+1)(N+1) + 16 bytes
    st = SimpleNamespace(pair_evals=12, stages=1)
    assert bench.bytes_per_launch(st, 4) == 24 * 3 * 5 + 16


def test_calibration_library_exports_its_microbenchmarks():
    if not os.path.exists(bench.CALIB_LIB):
        from pimcplusplus_amd import build
        build.build()
    lib = ctypes.CDLL(bench.CALIB_LIB)
    assert hasattr(lib, "pimc_calib_hbm") and hasattr(lib, "pimc_calib_lds")
