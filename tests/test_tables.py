"""Table ingest (.dm reader) and the Lennard-Jones generator."""

import os

import numpy as np
import pytest

from pimcplusplus_amd import tables
from tests.helpers import GOLDEN, REFERENCE


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree not present")
@pytest.mark.parametrize("name,src", [("testcoul_fit2.npz", "testcoul.dm"), ("ep_beta1_fit2.npz", "ep_beta1.0.dm")])
def test_read_dm_matches_fixture(name, src):
    t = tables.read_dm(os.path.join(REFERENCE, "PIMC++/examples", src))
    d = np.load(os.path.join(GOLDEN, name))
    assert np.array_equal(t.y, d["y"]) and np.array_equal(t.dy, d["dy"])


def test_testcoul_fixture_layout():
    d = np.load(os.path.join(GOLDEN, "testcoul_fit2.npz"))
    y = d["y"]
    # GRID 250 LOG .0025 10., 7 taus 1/64..1, fit 2 has 6 terms (+ potential column)
    assert y.shape == (7, 250, 7)
    assert int(d["grid_type"]) == tables.GRID_LOG
    assert float(d["grid_start"]) == 0.0025 and float(d["grid_end"]) == 10.0
    assert float(d["smallest_tau"]) == 0.015625
    assert np.all(y[:, -1, :] == 0.0)
    # column 0 is the potential on every tau level: e-p Coulomb, -1/r
    r = 0.0025 * (10.0 / 0.0025) ** (np.arange(250) / 249.0)
    assert np.allclose(y[0, :-1, 0], -1.0 / r[:-1], rtol=1e-4)
    assert all(np.array_equal(y[0, :, 0], y[i, :, 0]) for i in range(7))


def test_lj_table_diagonal_and_expansion():
    t = tables.lj_david_table(10.22, 2.556, 1.5, 7.0, 200, 0.025, 3)
    r = t.grid_points()
    for lev in range(3):
        tau = 0.025 * 2 ** lev
        assert np.allclose(t.y[lev, :-1, 1], tau * tables.lj_potential(r[:-1], 10.22, 2.556), rtol=1e-14)
    assert np.all(t.y[:, -1, :] == 0)
    # the fourth-order expansion agrees with the chord integral for small z, s (outside the core)
    tau = 0.05
    gl_x, gl_w = np.polynomial.legendre.leggauss(40)
    t_ = 0.5 * (gl_x + 1)
    w = 0.5 * gl_w
    for q in (3.0, 3.6, 4.5):
        f = tables._lj_f_derivs(np.array([q * q]), 10.22, 2.556)
        u = tables.straight_line_ukj(np.array([q]), tau, *f)[0]
        for s, z in ((0.1, 0.05), (0.2, -0.1), (0.15, 0.0)):
            rr, rp = q + z / 2, q - z / 2
            d = np.sqrt((1 - t_) * rr * rr + t_ * rp * rp - t_ * (1 - t_) * s * s)
            exact = tau * np.dot(w, tables.lj_potential(d, 10.22, 2.556)) - 0.5 * tau * (
                tables.lj_potential(rr, 10.22, 2.556) + tables.lj_potential(rp, 10.22, 2.556))
            approx = u[0] * s * s + u[1] * z * z + u[2] * s ** 4 + u[3] * z * z * s * s + u[4] * z ** 4
            assert abs(approx - exact) < 0.02 * abs(exact) + 1e-9
