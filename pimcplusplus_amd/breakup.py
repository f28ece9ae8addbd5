"""Optimized long-range breakup of a pair action (host-side table generation).

PIMC++ splits every long-ranged pair action U(r) into a short part, summed in real
space, and a long part W(r) summed over the k vectors |k| < kc
(LongRangeClass::OptimizedBreakup_U/_dU, PIMC++/src/Actions/LongRangeClass.cc:509-870,
after Natoli and Ceperley, J. Comput. Phys. 117, 171 (1995)).  Inside r_c = L/2 the
long part is a combination of locally piecewise quintic Hermite shape functions
(LPQHI, Common/Ewald/OptimizedBreakup.h:106-140); beyond r_c it equals U.  The
coefficients minimise, over k > kc, the weighted squared Fourier transform of W, so
that W is carried by the k < kc sum alone (OptimizedBreakup.cc:66-112, 196-320).

This module computes the same quantities with its own numerics:
  - the shape functions' Fourier transforms c_n(k) by Gauss-Legendre quadrature on
    each knot interval (the reference uses a closed-form recursion);
  - the transform of U beyond r_c from the fit U(r) ~ c1/r + c2/r^2 + c3/r^3 on
    [r_c, r_end] (PAcoulombBCFitClass::Setrc/Xk_U, Common/PairAction/PAcoulombBCFit.cc:467-548),
    with scipy's sine and cosine integrals;
  - the least-squares problem by an SVD with the reference's 1e-16 relative cutoff.

The tables it returns are what the library's pimc_replica_set_long_range and
pimc_replica_set_ulong take: Ulong_k, Ulong_r0, Ushort_k0, Ulong_k0 per level and
the quintic-spline data of Ulong(r) on LinearGrid(0, 0.75 |box|, 1000).  The same
routine with one "level", U = V = Z1Z2/r and the tail fixed to Z1Z2/r gives the
potential's tables for pimc_replica_set_vlong (LongRangeClass::OptimizedBreakup_V,
LongRangeClass.cc:979-1132).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# Rows alpha = 0, 1, 2 of the quintic Hermite shape functions in x = |r - r_i| / delta:
# value, slope and curvature at the knot (OptimizedBreakup.h:128-131).
_SHAPE = np.array([[1.0, 0.0, 0.0, -10.0, 15.0, -6.0],
                   [0.0, 1.0, 0.0, -6.0, 8.0, -3.0],
                   [0.0, 0.0, 0.5, -1.5, 1.5, -0.5]])

_GL_NODES, _GL_WEIGHTS = np.polynomial.legendre.leggauss(160)


class LPQHIBasis:
    """num_knots knots r_i = i delta on [0, rc] (delta = rc / (num_knots - 1)); element
    n = 3 i + alpha is shape function alpha of knot i, supported on [r_{i-1}, r_{i+1}]
    (LPQHI_BasisClass::h, OptimizedBreakup.cc:357-390)."""

    def __init__(self, rc: float, num_knots: int, box):
        if num_knots < 2:
            raise ValueError("num_knots must be at least 2")
        self.rc = float(rc)
        self.num_knots = int(num_knots)
        self.delta = self.rc / (self.num_knots - 1)
        self.omega = float(np.prod(box))

    @property
    def size(self) -> int:
        return 3 * self.num_knots

    def h(self, n: int, r) -> np.ndarray:
        r = np.asarray(r, dtype=float)
        i, alpha = divmod(n, 3)
        ra, rb, rc = self.delta * (i - 1), self.delta * i, min(self.rc, self.delta * (i + 1))
        out = np.zeros_like(r)
        left = (r > ra) & (r <= rb)
        right = (r > rb) & (r <= rc)
        xl = (rb - r[left]) / self.delta
        xr = (r[right] - rb) / self.delta
        out[left] = np.polyval(_SHAPE[alpha, ::-1], xl) * (-1.0) ** alpha
        out[right] = np.polyval(_SHAPE[alpha, ::-1], xr)
        return out

    def c(self, k: np.ndarray) -> np.ndarray:
        """c_n(k) = 4 pi / (Omega k) int_0^rc h_n(r) r sin(k r) dr for every element, [size][len(k)]
        (LPQHI_BasisClass::c, OptimizedBreakup.cc:393-421; BasisClass::c_numerical :40-52)."""
        k = np.asarray(k, dtype=float)
        out = np.zeros((self.size, k.size))
        # quadrature nodes on every knot interval [r_j, r_{j+1}]
        for j in range(self.num_knots - 1):
            a, b = self.delta * j, self.delta * (j + 1)
            r = 0.5 * (b - a) * _GL_NODES + 0.5 * (a + b)
            w = 0.5 * (b - a) * _GL_WEIGHTS
            sk = np.sin(np.outer(k, r)) * (w * r)[None, :]            # [nk][nodes]
            for n in (3 * j + np.arange(3)).tolist() + (3 * (j + 1) + np.arange(3)).tolist():
                out[n] += sk @ self.h(n, r)
        return out * (4.0 * math.pi / self.omega) / k[None, :]


def kpoints(box, kc: float, kcont: float, kmax: float, nshell: int = 4000):
    """(|k|, degeneracy) of the fit: every lattice k with kc < |k| < kcont, grouped by
    |k|, then nshell spherical shells from kcont to kmax with degeneracy = shell volume /
    reciprocal-cell volume (OptimizedBreakupClass::SetkVecs, OptimizedBreakup.cc:73-112)."""
    b = 2.0 * math.pi / np.asarray(box, dtype=float)
    m = [np.arange(-int(math.ceil(kcont / bd)), int(math.ceil(kcont / bd)) + 1) for bd in b]
    kx, ky, kz = np.meshgrid(m[0] * b[0], m[1] * b[1], m[2] * b[2], indexing="ij")
    k2 = (kx * kx + ky * ky + kz * kz).ravel()
    k2 = k2[(k2 > kc * kc) & (k2 < kcont * kcont)]
    kk = np.sqrt(k2)
    # group magnitudes that agree to 1e-12 (the reference merges |k1 - k2| <= 1e-12)
    kk.sort()
    starts = np.concatenate([[0], np.nonzero(np.diff(kk) > 1e-12)[0] + 1])
    mags = kk[starts]
    deg = np.diff(np.concatenate([starts, [kk.size]])).astype(float)
    dk = (kmax - kcont) / nshell
    k1 = kcont + dk * np.arange(nshell)
    k2s = k1 + dk
    shell = 0.5 * (k1 + k2s)
    sdeg = 4.0 * math.pi / 3.0 * (k2s ** 3 - k1 ** 3) / float(np.prod(b))
    return np.concatenate([mags, shell]), np.concatenate([deg, sdeg])


def fit_breakup(cnk: np.ndarray, weight: np.ndarray, target: np.ndarray):
    """t minimising sum_k weight_k (target_k - sum_n t_n c_n(k))^2 by the normal equations
    solved with an SVD whose singular values below 1e-16 of the largest are dropped
    (OptimizedBreakupClass::DoBreakup, OptimizedBreakup.cc:114-195).  Returns (t, chi2)."""
    A = (cnk * weight[None, :]) @ cnk.T
    b = cnk @ (weight * target)
    U, S, Vt = np.linalg.svd(A)
    Sinv = np.where(S < 1e-16 * S.max(), 0.0, 1.0 / np.where(S > 0, S, 1.0))
    t = Vt.T @ (Sinv * (U.T @ b))
    resid = target - t @ cnk
    return t, float(np.sum(weight * resid * resid))


def tail_fit(udiag, rc: float, r_end: float, num_points: int = 1000) -> np.ndarray:
    """Least-squares coefficients (c1, c2, c3) of U(r) ~ c1/r + c2/r^2 + c3/r^3 on
    [rc, r_end] (PAcoulombBCFitClass::Setrc, PAcoulombBCFit.cc:467-527)."""
    r = rc + np.arange(num_points) * ((r_end - rc) / (num_points - 1))
    F = np.stack([1.0 / r, 1.0 / r ** 2, 1.0 / r ** 3], axis=1)
    coef, *_ = np.linalg.lstsq(F, udiag(r), rcond=1e-15)
    return coef


def xk_tail(k: np.ndarray, rc: float, coef: np.ndarray) -> np.ndarray:
    """Xk(k) = -4 pi / k int_rc^inf r sin(k r) U(r) dr for U = c1/r + c2/r^2 + c3/r^3
    (PAcoulombBCFitClass::Xk_U, PAcoulombBCFit.cc:531-542): the three integrals in
    closed form with Si and Ci."""
    from scipy.special import sici
    k = np.asarray(k, dtype=float)
    kr = k * rc
    si, ci = sici(kr)
    c0 = -4.0 * math.pi / (k * k) * np.cos(kr)
    c1 = 4.0 * math.pi / k * (si - 0.5 * math.pi)
    c2 = 4.0 * math.pi / k * (k * ci - np.sin(kr) / rc)
    return coef[0] * c0 + coef[1] * c1 + coef[2] * c2


def _integrate(f, edges: np.ndarray) -> float:
    """Integral of the vectorised f over [edges[0], edges[-1]], Gauss-Legendre on each piece."""
    total = 0.0
    for a, b in zip(edges[:-1], edges[1:]):
        r = 0.5 * (b - a) * _GL_NODES + 0.5 * (a + b)
        total += float(np.dot(0.5 * (b - a) * _GL_WEIGHTS, f(r)))
    return total


@dataclass
class BreakupTables:
    ulong_k: np.ndarray      # [levels][nk]
    ulong_r0: np.ndarray     # [levels]
    ushort_k0: np.ndarray    # [levels]
    ulong_k0: np.ndarray     # [levels]
    r_grid: np.ndarray       # LongGrid(0, 0.75 |box|, 1000)
    ulong_r: np.ndarray      # [levels][1000] Ulong(r) on r_grid
    chi2: np.ndarray         # [levels]
    t: np.ndarray            # [levels][3 num_knots] shape-function coefficients
    basis: LPQHIBasis

    def ulong_in(self, level: int, r) -> np.ndarray:
        """The long part inside r_c, sum_n t_n h_n(r) (r <= r_c)."""
        r = np.asarray(r, dtype=float)
        return self.t[level] @ np.stack([self.basis.h(n, r) for n in range(self.basis.size)])


def optimized_breakup(udiag_levels, box, kc: float, kvecs: np.ndarray, num_knots: int = 20,
                      tail_end: float | None = None, tail_coef=None, short_of=None) -> BreakupTables:
    """The long-range tables of one pair from its diagonal action U(r, level).

    udiag_levels: callables U_l(r) (vectorised), one per tau level; each must extend past
    r_c = L/2 (a Coulomb tail).  tail_end: the end of the 1/r-polynomial fit (the pair
    table's q_end; default 2 r_c).  kvecs: the simulation's k vectors [nk][3].
    tail_coef: fixed (c1, c2, c3) instead of the fit (the potential's breakup uses
    (Z1Z2, 0, 0), PAcoulombBCFit.cc:511-520, 561-566).  short_of: the function whose
    difference with the long part is integrated for the short part's k = 0 transform, if
    not U itself (the dU breakup integrates V - dUlong, LongRangeClass.cc:181-186)."""
    box = np.asarray(box, dtype=float)
    omega = float(np.prod(box))
    rc = 0.5 * float(box.min())
    basis = LPQHIBasis(rc, num_knots, box)
    kavg = float(np.prod(2.0 * math.pi / box)) ** (1.0 / 3.0)
    kk, wk = kpoints(box, kc, 50.0 * kavg, 20.0 * math.pi / basis.delta)
    cnk = basis.c(kk)
    ksim = np.sqrt(np.einsum("ij,ij->i", kvecs, kvecs))
    csim = basis.c(ksim)
    rmax = 0.75 * math.sqrt(float(np.dot(box, box)))
    rg = np.arange(1000) * (rmax / 999.0)
    h_at = np.stack([basis.h(n, rg) for n in range(basis.size)])        # [size][1000]
    h0 = np.array([basis.h(n, np.array([0.0]))[0] for n in range(basis.size)])
    nl = len(udiag_levels)
    out = BreakupTables(np.zeros((nl, ksim.size)), np.zeros(nl), np.zeros(nl), np.zeros(nl), rg,
                        np.zeros((nl, rg.size)), np.zeros(nl), np.zeros((nl, basis.size)), basis)
    for lev, U in enumerate(udiag_levels):
        coef = np.asarray(tail_coef, dtype=float) if tail_coef is not None else \
            tail_fit(U, rc, tail_end if tail_end else 2.0 * rc)
        S = short_of if short_of is not None else U
        t, chi2 = fit_breakup(cnk, wk, xk_tail(kk, rc, coef) / omega)
        out.chi2[lev] = chi2
        out.t[lev] = t
        out.ulong_r0[lev] = float(t @ h0)
        inside = rg <= rc
        out.ulong_r[lev] = np.where(inside, t @ h_at, U(np.where(inside, rc, rg)))

        # k = 0 transforms of the short part on [0, rc] and of the long part on [0, rmax]
        # (LongRangeClass.cc:622-635): Gauss-Legendre on every knot interval and on 64
        # equal pieces of [rc, rmax]
        out.ushort_k0[lev] = 4.0 * math.pi / omega * _integrate(
            lambda r: r * r * (S(r) - t @ np.stack([basis.h(n, r) for n in range(basis.size)])),
            basis.delta * np.arange(num_knots))
        w_in = _integrate(lambda r: r * r * (t @ np.stack([basis.h(n, r) for n in range(basis.size)])),
                          basis.delta * np.arange(num_knots))
        w_out = _integrate(lambda r: r * r * U(r), np.linspace(rc, rmax, 65))
        out.ulong_k0[lev] = 4.0 * math.pi / omega * (w_in + w_out)
        out.ulong_k[lev] = t @ csim - xk_tail(ksim, rc, coef) / omega
    return out


def quintic_data(r_grid: np.ndarray, values: np.ndarray) -> np.ndarray:
    """QuinticSpline data {Y, B, C, D, E, F} per knot of the natural quintic interpolant of
    `values` on the uniform r_grid: [levels][n][6] (QuinticSpline::Init/operator(),
    Common/Splines/QuinticSpline.h:78-81, 204-218)."""
    from scipy.interpolate import make_interp_spline
    out = np.empty(values.shape + (6,))
    for lev in range(values.shape[0]):
        sp = make_interp_spline(r_grid, values[lev], k=5, bc_type=([(3, 0.0), (4, 0.0)], [(3, 0.0), (4, 0.0)]))
        out[lev] = np.stack([sp(r_grid, m) / math.factorial(m) for m in range(6)], axis=-1)
    return out
