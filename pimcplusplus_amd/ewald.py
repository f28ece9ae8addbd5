"""Long-range (k-space) tables of a Coulomb pair in the classic Ewald split.

The pair potential Z1 Z2 / r splits into erfc(a r)/r (short) and erf(a r)/r
(long).  In LongRangeClass::SingleAction's conventions (LongRangeClass.cc:313-367,
half-space k set, the homologous sum weighted 0.5 * 2.0) the tables per level
with tau_l = tau 2^l are

    Ulong_k   = tau_l Z1Z2 4 pi exp(-k^2 / 4 a^2) / (Omega k^2)
    Ulong_r0  = tau_l Z1Z2 2 a / sqrt(pi)          (the long part at r = 0)
    Ushort_k0 = tau_l Z1Z2 pi / (a^2 Omega)         (k = 0 of the short part)
    Ulong_k0  = 0                                   (divergent; neutral systems cancel it)

These are the closed-form tables of the plain Ewald breakup; PIMC++ fits an
optimized breakup instead (OptimizedBreakup.cc), which is a setup-time step.
"""

from __future__ import annotations

import math

import numpy as np


def ewald_tables(kvecs: np.ndarray, box, z1z2: float, alpha: float, tau: float, num_levels: int):
    """Returns (ulong_k [levels][nk], ulong_r0 [levels], ushort_k0 [levels], ulong_k0 [levels])."""
    omega = float(np.prod(box))
    k2 = np.einsum("ij,ij->i", kvecs, kvecs)
    vk = z1z2 * 4.0 * math.pi * np.exp(-k2 / (4.0 * alpha * alpha)) / (omega * k2)
    taus = tau * 2.0 ** np.arange(num_levels)
    ulong_k = taus[:, None] * vk[None, :]
    ulong_r0 = taus * z1z2 * 2.0 * alpha / math.sqrt(math.pi)
    ushort_k0 = taus * z1z2 * math.pi / (alpha * alpha * omega)
    ulong_k0 = np.zeros(num_levels)
    return ulong_k, ulong_r0, ushort_k0, ulong_k0


def ewald_dbeta_tables(kvecs: np.ndarray, box, z1z2: float, alpha: float, tau: float, num_levels: int):
    """The beta derivatives (dUlong_k, dUlong_r0, dUshort_k0, dUlong_k0): every table above is
    tau_l times a potential, so its derivative is that potential at every level."""
    ulong_k, ulong_r0, ushort_k0, ulong_k0 = ewald_tables(kvecs, box, z1z2, alpha, tau, num_levels)
    taus = tau * 2.0 ** np.arange(num_levels)
    return ulong_k / taus[:, None], ulong_r0 / taus, ushort_k0 / taus, ulong_k0 / taus


def kset_size(n: int) -> int:
    """Half-space k vectors with |k| < n 2 pi / L in a cubic box: (#{m in Z^3 : 0 < |m|^2 < n^2}) / 2."""
    r = np.arange(-n, n + 1)
    m2 = r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2
    return int(((m2 > 0) & (m2 < n * n)).sum() // 2)


def ulong_quintic(z1z2: float, alpha: float, tau: float, num_levels: int, r_end: float, n: int = 1000):
    """Ulong(level)(r) = tau_l Z1Z2 erf(alpha r)/r as the reference's QuinticSpline data on
    LinearGrid(0, r_end, n): coefficients [levels][n][6] = {Y, B, C, D, E, F}, the Taylor
    coefficients f^(m)(x_i)/m! of the natural quintic interpolant (third and fourth
    derivatives zero at both ends) at every knot."""
    from scipy.interpolate import make_interp_spline
    from scipy.special import erf
    x = 0.0 + np.arange(n) * ((r_end - 0.0) / float(n - 1))
    with np.errstate(invalid="ignore", divide="ignore"):
        f = np.where(x > 0, erf(alpha * x) / np.where(x > 0, x, 1.0), 2.0 * alpha / math.sqrt(math.pi))
    sp = make_interp_spline(x, f, k=5, bc_type=([(3, 0.0), (4, 0.0)], [(3, 0.0), (4, 0.0)]))
    base = np.stack([sp(x, m) / math.factorial(m) for m in range(6)], axis=-1)   # [n][6]
    out = np.empty((num_levels, n, 6))
    for lev in range(num_levels):
        out[lev] = tau * 2 ** lev * z1z2 * base
    return 0.0, float(x[-1]), out


def dulong_quintic(z1z2: float, alpha: float, tau: float, num_levels: int, r_end: float, n: int = 1000):
    """dUlong(level)(r) = Z1Z2 erf(alpha r)/r on the grid of ulong_quintic (the derivative of
    tau_l Vlong with respect to tau_l)."""
    start, end, coefs = ulong_quintic(z1z2, alpha, tau, num_levels, r_end, n)
    taus = tau * 2.0 ** np.arange(num_levels)
    return start, end, coefs / taus[:, None, None]
