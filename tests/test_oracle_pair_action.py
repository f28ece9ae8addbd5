"""Pins DavidPAClass::U of the oracle (DavidPAClass.cc:92-102, 468-511).

Column mapping and the expansion u_kj z^{2j} s^{2(k-j)} are pinned with one-hot
tables; spline values with scipy; the special branches with closed forms; the
shipped squarer table (testcoul.dm) with a scipy-spline evaluation."""

import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from tests.helpers import GOLDEN, grid_points, small_random_table


def _fit(lib, y, gtype=0, start=0.5, end=4.0, nt=2, tau0=0.1, desired=None):
    return lib.fit_create_david(gtype, start, end, y, nt, tau0, tau0 if desired is None else desired)


@pytest.mark.parametrize("k,j", [(1, 0), (1, 1), (2, 0), (2, 1), (2, 2)])
def test_off_diagonal_column_mapping(oracle, k, j):
    # one-hot: only column k(k+1)/2 + j + 1 is a nonzero constant -> U = c z^{2j} s^{2(k-j)}
    n = 20
    y = np.zeros((1, n, 7))
    col = k * (k + 1) // 2 + j + 1
    y[0, :, col] = 0.7
    y[0, -1, :] = 0.0  # the reader's convention; keep the interior flat
    y[0, -1, col] = 0.7
    fid = _fit(oracle, y, start=1.0, end=5.0)
    rng = np.random.default_rng(4)
    q = rng.uniform(2.0, 3.0, 50)
    s = rng.uniform(0.1, 0.9, 50)
    z = rng.uniform(-1, 1, 50) * s
    u = oracle.fit_eval_u(fid, 0, q, z, s * s)
    want = 0.7 * z ** (2 * j) * s ** (2 * (k - j))
    assert np.allclose(u, want, rtol=1e-12, atol=1e-15)


def test_diagonal_branches(oracle):
    rng = np.random.default_rng(5)
    y = small_random_table(rng, num_points=40, num_tau=1)
    fid = _fit(oracle, y, start=0.5, end=4.0)
    x = grid_points(0, 0.5, 4.0, 40)
    u00 = CubicSpline(x, y[0, :, 1], bc_type=((2, 0.0), (1, 0.0)))
    q = rng.uniform(1.0, 3.0, 200)
    z = rng.uniform(-0.5, 0.5, 200)
    # s = 0: diagonal only, (u00(r) + u00(r'))/2
    u = oracle.fit_eval_u(fid, 0, q, z, np.zeros(200))
    assert np.allclose(u, 0.5 * (u00(q + z / 2) + u00(q - z / 2)), rtol=0, atol=1e-12)
    # r or r' below rmin: u00(rmin), whatever q, z, s
    u = oracle.fit_eval_u(fid, 0, np.full(3, 0.6), np.array([0.5, -0.5, 0.3]), np.full(3, 0.2))
    assert np.allclose(u, y[0, 0, 1], rtol=0, atol=0)
    # q at or beyond the grid end: r, r' clamp to End and no off-diagonal term
    u = oracle.fit_eval_u(fid, 0, np.array([4.0, 5.0]), np.array([0.0, 0.2]), np.array([0.3, 0.3]))
    assert np.allclose(u, [u00(4.0), u00(4.0)], atol=1e-13)


def test_symmetric_in_z(oracle):
    rng = np.random.default_rng(6)
    y = small_random_table(rng, num_points=50, num_tau=2)
    fid = _fit(oracle, y, start=0.5, end=4.0, tau0=0.1, desired=0.1)
    q = rng.uniform(1.0, 3.0, 300)
    s = rng.uniform(0.01, 1.0, 300)
    z = rng.uniform(-1, 1, 300) * s
    for lev in (0, 1):
        a = oracle.fit_eval_u(fid, lev, q, z, s * s)
        b = oracle.fit_eval_u(fid, lev, q, -z, s * s)
        assert np.allclose(a, b, rtol=1e-14, atol=1e-15)


def test_tau_level_selection(oracle):
    # level + TauPos picks the table: tau_i = tau_min 2^i, TauPos where |tau_i - tau| < 1e-4
    rng = np.random.default_rng(7)
    y = small_random_table(rng, num_points=30, num_tau=4)
    f_lo = _fit(oracle, y, tau0=0.025, desired=0.025)     # TauPos 0
    f_hi = _fit(oracle, y, tau0=0.025, desired=0.05)      # TauPos 1
    q = rng.uniform(1.0, 3.0, 40)
    z = np.zeros(40)
    s2 = np.full(40, 0.04)
    assert np.array_equal(oracle.fit_eval_u(f_lo, 1, q, z, s2), oracle.fit_eval_u(f_hi, 0, q, z, s2))
    assert np.array_equal(oracle.fit_eval_u(f_lo, 3, q, z, s2), oracle.fit_eval_u(f_hi, 2, q, z, s2))


def scipy_david_u(y, x, q, z, s2, nt):
    """calcUsqzFast over scipy splines (natural start, zero end slope)."""
    sp = [CubicSpline(x, y[:, c], bc_type=((2, 0.0), (1, 0.0))) for c in range(y.shape[1])]
    rmin, end = x[0], x[-1]
    r = np.minimum(q + 0.5 * z, end)
    rp = np.minimum(q - 0.5 * z, end)
    out = 0.5 * (sp[1](np.maximum(r, rmin)) + sp[1](np.maximum(rp, rmin)))
    low = (r < rmin) | (rp < rmin)
    out[low] = sp[1](rmin)
    s = np.sqrt(s2)
    use = (~low) & (s > 0) & (q < end)
    z2 = z * z
    for k in range(1, nt + 1):
        for j in range(0, k + 1):
            c = k * (k + 1) // 2 + j + 1
            out[use] += sp[c](q[use]) * z2[use] ** j * s2[use] ** (k - j)
    return out


def test_testcoul_table_against_scipy(oracle):
    d = np.load(os.path.join(GOLDEN, "testcoul_fit2.npz"))
    y = d["y"]
    # tau = 1/64 * 2^3 = 0.125
    fid = oracle.fit_create_david(int(d["grid_type"]), float(d["grid_start"]), float(d["grid_end"]), y, 2,
                                  float(d["smallest_tau"]), 0.125)
    x = grid_points(1, float(d["grid_start"]), float(d["grid_end"]), y.shape[1])
    rng = np.random.default_rng(8)
    q = np.exp(rng.uniform(np.log(0.01), np.log(9.9), 3000))
    s = rng.uniform(0, 1, 3000) * np.minimum(2 * q, 2.0)
    z = rng.uniform(-1, 1, 3000) * s
    for lev in (0, 1, 2):
        u = oracle.fit_eval_u(fid, lev, q, z, s * s)
        ref = scipy_david_u(y[3 + lev], x, q, z, s * s, 2)
        assert np.max(np.abs(u - ref) / (1 + np.abs(ref))) < 1e-11
