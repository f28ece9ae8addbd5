"""Pins the oracle's grids and MultiCubicSpline against closed forms and scipy."""

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from tests.helpers import grid_points, multicubic, reverse_map

LIN, LOG = 0, 1


def test_linear_grid_points_and_reverse_map():
    x = grid_points(LIN, 0.5, 7.15, 200)
    delta = (7.15 - 0.5) / 199.0
    assert np.array_equal(x, 0.5 + np.arange(200) * delta)
    rng = np.random.default_rng(1)
    r = rng.uniform(0.5, 7.15, 200000)
    got = reverse_map(LIN, 0.5, 7.15, 200, r, literal=True)
    brute = np.searchsorted(x, r, side="right") - 1
    diff = got != brute
    # they differ only within rounding distance of a knot
    assert np.all(np.min(np.abs(r[diff, None] - x[None, :]), axis=1) < 1e-12 * 8)
    assert diff.sum() < 10


def test_linear_reverse_map_at_knots_rounds_half_to_even():
    x = grid_points(LIN, 0.0, 10.0, 11)
    got = reverse_map(LIN, 0.0, 10.0, 11, x[1:10], literal=True)
    # nearbyint(i - 0.5) with ties to even: i-1 when i-1 is even, else i
    want = np.array([i - 1 if (i - 1) % 2 == 0 else i for i in range(1, 10)])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("start,end,n", [(0.0025, 10.0, 250), (0.1, 12.0, 64)])
def test_log_grid_corrected_reverse_map_is_the_true_interval(start, end, n):
    x = grid_points(LOG, start, end, n)
    spacing = (end / start) ** (1.0 / (n - 1.0))
    assert np.allclose(x, start * spacing ** np.arange(n), rtol=1e-14, atol=0)
    rng = np.random.default_rng(2)
    r = np.concatenate([np.exp(rng.uniform(np.log(start), np.log(x[-1]), 100000)), x[1:-1],
                        np.nextafter(x[1:-1], 0), np.nextafter(x[1:-1], np.inf)])
    got = reverse_map(LOG, start, end, n, r)
    brute = np.searchsorted(x, r, side="right") - 1
    assert np.array_equal(got, brute)
    lit = reverse_map(LOG, start, end, n, r, literal=True)
    d = lit != got
    # the reference's literal log formula disagrees only at knot neighbours
    assert np.all(np.min(np.abs(r[d, None] / x[None, :] - 1.0), axis=1) < 1e-13)


@pytest.mark.parametrize("gtype,start,end", [(LIN, 0.3, 5.0), (LOG, 0.01, 8.0)])
def test_multicubic_matches_scipy_natural_start_clamped_end(gtype, start, end):
    n, nc = 60, 4
    x = grid_points(gtype, start, end, n)
    rng = np.random.default_rng(3)
    y = np.stack([np.sin((c + 1) * x) * np.exp(-0.2 * x) + 0.3 * c for c in range(nc)], axis=1)
    y += 0.01 * rng.normal(size=y.shape)
    sd = np.full(nc, 5.0e30)     # natural
    ed = np.zeros(nc)            # first derivative 0 (the .dm reader's end condition)
    xs = np.concatenate([rng.uniform(x[0], x[-1], 5000), x])
    d2, vals = multicubic(gtype, start, end, y, sd, ed, xs)
    for c in range(nc):
        ref = CubicSpline(x, y[:, c], bc_type=((2, 0.0), (1, 0.0)))
        scale = np.max(np.abs(y[:, c]))
        assert np.max(np.abs(vals[:, c] - ref(xs))) < 1e-12 * scale
        assert np.max(np.abs(d2[:, c] - ref(x, 2))) < 1e-9 * np.max(np.abs(d2[:, c]))
    # interpolation at the knots
    assert np.max(np.abs(vals[-n:, :] - y)) < 1e-13 * np.max(np.abs(y))


def test_multicubic_natural_reproduces_linear_function():
    x = grid_points(LIN, 1.0, 3.0, 17)
    y = (2.5 - 0.75 * x)[:, None]
    xs = np.linspace(1.0, 3.0, 1001)
    d2, vals = multicubic(LIN, 1.0, 3.0, y, [5e30], [5e30], xs, col=0)
    assert np.max(np.abs(d2)) < 1e-13
    assert np.max(np.abs(vals - (2.5 - 0.75 * xs))) < 1e-14


def test_multicubic_clamped_start_matches_scipy():
    x = grid_points(LIN, 0.0, 2.0, 30)
    y = np.cos(3 * x)[:, None]
    xs = np.linspace(0, 2, 777)
    _, vals = multicubic(LIN, 0.0, 2.0, y, [-0.4], [5e30], xs, col=0)
    ref = CubicSpline(x, y[:, 0], bc_type=((1, -0.4), (2, 0.0)))
    assert np.max(np.abs(vals - ref(xs))) < 1e-13
