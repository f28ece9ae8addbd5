"""Pair-action tables in the layout of PIMC++'s DavidPAClass.

Two sources:

* `read_dm` parses a squarer ``.dm`` file as DavidPAClass::ReadDavidSquarerFile
  does (Common/PairAction/DavidPAClass.cc:639-912): the potential block, then
  the rank-3 ``fit= n`` tables; the table whose fit order NMax is 2 becomes the
  spline data, with column 0 the potential and the last grid row zeroed.
* `read_squarer_h5` reads the HDF5 form of a squarer file as
  DavidPAClass::ReadDavidSquarerFileHDF5 does (DavidPAClass.cc:932-1224), and
  `read_coulomb_bc_h5` a ``coulombBCfit`` file as ReadPAFit and
  PAcoulombBCFitClass::Read do (PAFit.h:30-69, PAcoulombBCFit.cc:242-311).
* `lj_david_table` builds the same layout for a Lennard-Jones pair.  PIMC++
  ships no Lennard-Jones pair action, so this is this package's own generator:
  the diagonal is the primitive end-point action u00(r) = tau V(r) and the
  off-diagonal u_kj(q) are the fourth-order expansion, in the DavidPAClass
  basis z^{2j} s^{2(k-j)}, of the straight-line (semiclassical) action
  tau * int_0^1 V(|r + t(r' - r)|) dt minus the end-point average.
"""

from __future__ import annotations

import re
from dataclasses import dataclass

import numpy as np

GRID_LINEAR, GRID_LOG = 0, 1


@dataclass
class DavidTable:
    grid_type: int
    grid_start: float
    grid_end: float
    y: np.ndarray            # [num_tau][num_points][num_cols]
    num_terms: int
    smallest_tau: float
    dy: np.ndarray | None = None   # beta-derivative table, same layout (dukj), if known

    @property
    def num_tau(self) -> int:
        return self.y.shape[0]

    @property
    def num_points(self) -> int:
        return self.y.shape[1]

    def grid_points(self) -> np.ndarray:
        n = self.num_points
        if self.grid_type == GRID_LINEAR:
            delta = (self.grid_end - self.grid_start) / float(n - 1)
            return self.grid_start + np.arange(n, dtype=np.float64) * delta
        spacing = (self.grid_end / self.grid_start) ** (1.0 / (n - 1.0))
        return self.grid_start * spacing ** np.arange(n, dtype=np.float64)


# --------------------------------------------------------------------------
# .dm squarer files
# --------------------------------------------------------------------------
_NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[EeDd][-+]?\d+)?")


def _floats(line: str) -> list[float]:
    return [float(t.replace("D", "E").replace("d", "e")) for t in _NUM.findall(line)]


def read_dm(path: str) -> DavidTable:
    """Reads a squarer .dm file into the DavidPAClass table layout.

    Follows DavidPAClass::ReadDavidSquarerFile (DavidPAClass.cc:639-912): the
    number of fits is the fourth number on the SQUARER line; the potential is
    the first RANK 2 block; each RANK 3 block carries a GRID 1 (r grid), a
    GRID 3 LOG tau grid and a 'BEGIN density matrix table <NMax> <derv>' header
    followed by the table in Fortran order (r fastest, then term, then tau).
    The NMax == 2 tables (U and its beta derivative) become the spline data.
    """
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0

    def skip_to(key: str) -> str:
        nonlocal i
        while i < len(lines) and key not in lines[i]:
            i += 1
        if i >= len(lines):
            raise ValueError(f"{path}: no '{key}' found")
        line = lines[i]
        i += 1
        return line

    sq = skip_to("SQUARER").split()
    num_fits = int(float(sq[4]))
    rank_line = skip_to("RANK")
    npot = int(rank_line.split()[2])
    skip_to("potential")
    potential: list[float] = []
    while len(potential) < npot:
        potential += _floats(lines[i])
        i += 1
    potential_arr = np.array(potential[:npot])

    def read_tables(expect_derv: int):
        nonlocal i
        found = None
        counter = 0
        while counter <= num_fits:
            words = skip_to("RANK").split()
            if int(words[1]) != 3:
                continue
            ngrid, nukj, ntau = int(words[2]), int(words[3]), int(words[4])
            g = skip_to("GRID 1").split()
            gtype, gstart, gend = g[2], float(g[3]), float(g[4])
            t = skip_to("GRID   3").split()
            if t[2] != "LOG":
                raise ValueError(f"{path}: tau grid is not LOG")
            tmin, tmax = float(t[3]), float(t[4])
            begin = skip_to("BEGIN density matrix table")
            ints = [int(v) for v in re.findall(r"=\s*(\d+)", begin)]
            nmax, derv = ints[0], ints[1]
            if derv != expect_derv:
                raise ValueError(f"{path}: table derv {derv}, expected {expect_derv}")
            vals: list[float] = []
            need = ngrid * nukj * ntau
            while len(vals) < need:
                vals += _floats(lines[i])
                i += 1
            tab = np.array(vals[:need]).reshape(ntau, nukj, ngrid)  # Fortran order
            if nmax == 2:
                found = (gtype, gstart, gend, ngrid, nukj, ntau, tmin, tmax, tab)
            counter += 1
        if found is None:
            raise ValueError(f"{path}: no NMax == 2 table")
        return found

    gtype, gstart, gend, ngrid, nukj, ntau, tmin, tmax, u = read_tables(1)
    _, _, _, _, _, _, _, _, du = read_tables(2)

    def assemble(tab):
        y = np.zeros((ntau, ngrid, nukj + 1))
        y[:, :, 0] = potential_arr[None, :]
        y[:, :, 1:] = np.transpose(tab, (0, 2, 1))
        y[:, ngrid - 1, :] = 0.0        # DavidPAClass.cc:759,871
        return y

    grid_type = {"LINEAR": GRID_LINEAR, "LOG": GRID_LOG}[gtype]
    # fit order 2 -> NumOffDiagonalTerms 2 uses columns up to 2*3/2+2+1 = 6
    return DavidTable(grid_type=grid_type, grid_start=gstart, grid_end=gend, y=assemble(u),
                      num_terms=2, smallest_tau=tmin, dy=assemble(du))


# --------------------------------------------------------------------------
# HDF5 squarer files (PIMC++'s IO layer; see pimcplusplus_amd.h5)
# --------------------------------------------------------------------------
def read_squarer_h5(path: str) -> DavidTable:
    """Reads the HDF5 form of a squarer file into the DavidPAClass table layout.

    Follows DavidPAClass::ReadDavidSquarerFileHDF5 (DavidPAClass.cc:932-1224):
    Squarer/NumFits; Potential/Data (column 0 of every level); sections Ukj<c> and
    dUkjdBeta<c> for c = 0..NumFits, each with Rank (3), NumUkj, NumTau, a Grid section
    (NumGridPoints, Type LINEAR or LOG, Start, End), Taus, NMax, Derv (1 for U, 2 for
    dU/dbeta) and Data [grid][ukj][tau].  The NMax == 2 tables become the spline data
    with the last grid row zeroed (:1045, :1170), as in the .dm reader.
    """
    import warnings

    from . import h5
    with h5.open_file(path) as f:
        with f.group("Squarer") as sq:
            num_fits = int(sq.read("NumFits"))
        with f.group("Potential") as pg:
            potential = np.atleast_1d(np.asarray(pg.read("Data"), dtype=np.float64))

        def tables(prefix: str, expect_derv: int):
            found = None
            for counter in range(num_fits + 1):
                with f.group(f"{prefix}{counter}") as g:
                    if int(g.read("Rank")) != 3:
                        raise ValueError(f"{path}: section {prefix}{counter} is not a rank-3 table")
                    nukj, ntau = int(g.read("NumUkj")), int(g.read("NumTau"))
                    with g.group("Grid") as gg:
                        ngrid = int(gg.read("NumGridPoints"))
                        gtype, gstart, gend = gg.read("Type"), float(gg.read("Start")), float(gg.read("End"))
                    taus = np.atleast_1d(np.asarray(g.read("Taus"), dtype=np.float64))
                    nmax, derv = int(g.read("NMax")), int(g.read("Derv"))
                    data = np.asarray(g.read("Data"), dtype=np.float64)
                if data.shape != (ngrid, nukj, ntau):
                    raise ValueError(f"{path}: {prefix}{counter}/Data has shape {data.shape}, "
                                     f"expected {(ngrid, nukj, ntau)}")
                if gtype not in ("LINEAR", "LOG"):
                    raise ValueError(f"{path}: unrecognized grid type {gtype!r} in {prefix}{counter}")
                if derv != expect_derv:   # the reference reports this and goes on (:1030-1032, :1127-1129)
                    warnings.warn(f"{path}: {prefix}{counter} has Derv {derv}, expected {expect_derv}")
                num_tau_calc = int(np.floor(np.log(taus[-1] / taus[0]) / np.log(2.0) + 0.5 + 1.0))
                if num_tau_calc != ntau:   # reported by the reference too (:1021-1026)
                    warnings.warn(f"{path}: {prefix}{counter}: NumTau {ntau} against {num_tau_calc} from Taus")
                if nmax == 2:
                    found = (gtype, gstart, gend, ngrid, nukj, ntau, float(taus[0]), data)
            if found is None:
                raise ValueError(f"{path}: no NMax == 2 table among the {prefix} sections")
            return found

        gtype, gstart, gend, ngrid, nukj, ntau, tmin, u = tables("Ukj", 1)
        _, _, _, ngrid_d, nukj_d, ntau_d, _, du = tables("dUkjdBeta", 2)
    if potential.shape != (ngrid,) or (ngrid_d, nukj_d, ntau_d) != (ngrid, nukj, ntau):
        raise ValueError(f"{path}: the potential, U and dU/dbeta tables disagree in size")

    def assemble(tab):   # [grid][ukj][tau] -> [tau][grid][col], column 0 the potential
        y = np.zeros((ntau, ngrid, nukj + 1))
        y[:, :, 0] = potential[None, :]
        y[:, :, 1:] = np.transpose(tab, (2, 0, 1))
        y[:, ngrid - 1, :] = 0.0
        return y

    grid_type = GRID_LINEAR if gtype == "LINEAR" else GRID_LOG
    return DavidTable(grid_type=grid_type, grid_start=gstart, grid_end=gend, y=assemble(u), num_terms=2,
                      smallest_tau=tmin, dy=assemble(du))


# --------------------------------------------------------------------------
# Lennard-Jones generator
# --------------------------------------------------------------------------
def load_npz(path: str) -> DavidTable:
    """A DavidTable saved with numpy.savez (fields y, dy, grid_type, grid_start, grid_end,
    smallest_tau, num_terms)."""
    d = np.load(path, allow_pickle=False)
    return DavidTable(int(d["grid_type"]), float(d["grid_start"]), float(d["grid_end"]), d["y"],
                      int(d["num_terms"]), float(d["smallest_tau"]), d["dy"])


def lj_potential(r, eps: float, sigma: float):
    sr6 = (sigma / np.asarray(r, dtype=np.float64)) ** 6
    return 4.0 * eps * (sr6 * sr6 - sr6)


def num_cols_for_terms(num_terms: int) -> int:
    return num_terms * (num_terms + 1) // 2 + num_terms + 2


def _lj_f_derivs(x, eps: float, sigma: float):
    """f(x) = V(sqrt(x)) for Lennard-Jones and its first four x-derivatives."""
    s6, s12 = sigma ** 6, sigma ** 12
    c = 4.0 * eps
    f1 = c * (-6.0 * s12 * x ** -7 + 3.0 * s6 * x ** -4)
    f2 = c * (42.0 * s12 * x ** -8 - 12.0 * s6 * x ** -5)
    f3 = c * (-336.0 * s12 * x ** -9 + 60.0 * s6 * x ** -6)
    f4 = c * (3024.0 * s12 * x ** -10 - 360.0 * s6 * x ** -7)
    return f1, f2, f3, f4


def straight_line_ukj(q, tau: float, f1, f2, f3, f4):
    """Off-diagonal coefficients of tau*int_0^1 V(|r + t(r'-r)|)dt - tau(V(r)+V(r'))/2.

    With r = q + z/2, r' = q - z/2, s = |r - r'| and f(x) = V(sqrt x), the squared
    distance along the chord is q^2 + z^2/4 + (1-2t) q z - t(1-t) s^2; expanding both
    the chord integral and the end-point average to fourth order in (z, s) gives
    u10 s^2 + u11 z^2 + u20 s^4 + u21 z^2 s^2 + u22 z^4 with the coefficients below
    (derivatives of f taken at x = q^2).
    """
    q2 = q * q
    u10 = -tau * f1 / 6.0
    u11 = -tau * f2 * q2 / 3.0
    u20 = tau * f2 / 60.0
    u21 = -tau * f2 / 24.0 - tau * f3 * q2 / 60.0
    u22 = -tau * f3 * q2 / 12.0 - tau * f4 * q2 * q2 / 30.0
    return np.stack([u10, u11, u20, u21, u22], axis=-1)


def lj_david_table(eps: float, sigma: float, r_min: float, r_max: float, num_points: int,
                   smallest_tau: float, num_tau: int, num_terms: int = 2,
                   damp: tuple = (0.80, 0.90)) -> DavidTable:
    """A Lennard-Jones pair action table on LinearGrid(r_min, r_max, num_points).

    Column 0 is V, column 1 u00 = tau V, columns 2..6 the straight-line expansion
    coefficients (straight_line_ukj), switched off smoothly inside the repulsive core
    (q between damp[0]*sigma and damp[1]*sigma) where the expansion does not converge.
    num_terms is 0 (primitive action) or 2.
    """
    if num_terms not in (0, 2):
        raise ValueError("num_terms must be 0 or 2")
    delta = (r_max - r_min) / float(num_points - 1)
    r = r_min + np.arange(num_points) * delta
    vr = lj_potential(r, eps, sigma)
    ncols = num_cols_for_terms(num_terms)
    y = np.zeros((num_tau, num_points, ncols))
    dy = np.zeros_like(y)
    a, b = damp[0] * sigma, damp[1] * sigma
    t = np.clip((r - a) / (b - a), 0.0, 1.0)
    switch = t * t * (3.0 - 2.0 * t)
    f1, f2, f3, f4 = _lj_f_derivs(r * r, eps, sigma)
    for lev in range(num_tau):
        tau = smallest_tau * 2.0 ** lev
        y[lev, :, 0] = vr
        y[lev, :, 1] = tau * vr
        if num_terms == 2:
            y[lev, :, 2:7] = switch[:, None] * straight_line_ukj(r, tau, f1, f2, f3, f4)
        # the straight-line action is linear in tau: dU/dtau = U/tau; column 0 stays V.  The
        # diagonal column holds d(u00)/dtau - V, the .dm convention that calcUsqz
        # compensates by adding V back (DavidPAClass.cc:362-366): here 0.
        dy[lev, :, 0] = vr
        dy[lev, :, 1:] = y[lev, :, 1:] / tau
        dy[lev, :, 1] = 0.0
    y[:, num_points - 1, :] = 0.0    # the .dm reader's convention (DavidPAClass.cc:759)
    dy[:, num_points - 1, :] = 0.0
    return DavidTable(grid_type=GRID_LINEAR, grid_start=float(r_min), grid_end=float(r_max), y=y,
                      num_terms=num_terms, smallest_tau=smallest_tau, dy=dy)


# --------------------------------------------------------------------------
# Coulomb tables in the PAcoulombBCFitClass layout: U over (q, t = s / 2q)
# --------------------------------------------------------------------------
@dataclass
class CoulombBCTable:
    q_grid: tuple              # (grid_type, start, end)
    t_grid: tuple
    umat: np.ndarray           # [num_betas][nq][nt]
    smallest_beta: float
    z1z2: float
    dumat: np.ndarray | None = None   # dU/dbeta, same layout

    def grid_points(self, which: str) -> np.ndarray:
        gtype, start, end = self.q_grid if which == "q" else self.t_grid
        n = self.umat.shape[1] if which == "q" else self.umat.shape[2]
        return _grid(gtype, start, end, n)

    def udiag(self, level: int, derivative: bool = False):
        """The diagonal action U(r, r; beta_l) (or dU/dbeta) as a vectorised function of r
        (PAcoulombBCFitClass::Udiag / dUdiag, Common/PairAction/PAcoulombBCFit.cc:384-438): the
        bicubic at t = 0, which is the natural cubic spline through the t = 0 column along q,
        inside the q grid; beyond it the Coulomb tail beta_l Z1Z2 / r (Z1Z2 / r for dU) plus
        the end mismatch decaying as (r_end / r)^3."""
        from scipy.interpolate import CubicSpline
        q = self.grid_points("q")
        table = self.dumat if derivative else self.umat
        if table is None:
            raise ValueError("this table has no dU/dbeta data")
        sp = CubicSpline(q, table[level][:, 0], bc_type="natural")
        rend = q[-1]
        scale = 1.0 if derivative else self.smallest_beta * 2.0 ** level
        delta = float(sp(rend)) - scale * self.z1z2 / rend

        def f(r):
            r = np.asarray(r, dtype=float)
            inside = r <= rend * 1.0000001
            rr = np.where(inside, rend, r)
            return np.where(inside, sp(np.minimum(r, rend)), scale * self.z1z2 / rr + delta * rend ** 3 / rr ** 3)
        return f


def _grid(gtype: int, start: float, end: float, n: int) -> np.ndarray:
    if gtype == GRID_LINEAR:
        delta = (end - start) / float(n - 1)
        return start + np.arange(n, dtype=np.float64) * delta
    spacing = (end / start) ** (1.0 / (n - 1.0))
    return start * spacing ** np.arange(n, dtype=np.float64)


def david_u(table: DavidTable, level_index: int, q, z, s2):
    """DavidPAClass::U of a DavidTable level (natural start, zero end slope), evaluated
    with scipy splines; used to convert squarer tables."""
    from scipy.interpolate import CubicSpline
    x = table.grid_points()
    y = table.y[level_index]
    sp = [CubicSpline(x, y[:, c], bc_type=((2, 0.0), (1, 0.0))) for c in range(y.shape[1])]
    q, z, s2 = (np.asarray(v, dtype=np.float64) for v in (q, z, s2))
    rmin, end = x[0], x[-1]
    r = np.minimum(q + 0.5 * z, end)
    rp = np.minimum(q - 0.5 * z, end)
    out = 0.5 * (sp[1](np.maximum(r, rmin)) + sp[1](np.maximum(rp, rmin)))
    low = (r < rmin) | (rp < rmin)
    out = np.where(low, sp[1](rmin), out)
    use = (~low) & (s2 > 0) & (q < end)
    z2 = z * z
    for k in range(1, table.num_terms + 1):
        for j in range(0, k + 1):
            c = k * (k + 1) // 2 + j + 1
            out = out + np.where(use, sp[c](np.minimum(q, end)) * z2 ** j * s2 ** (k - j), 0.0)
    return out


def david_du(table: DavidTable, level_index: int, q, z, s2):
    """DavidPAClass::dU (calcUsqz, DavidPAClass.cc:291-463) of a DavidTable level with its
    beta-derivative table, evaluated with scipy splines: r and r' clamped into the grid,
    dU = (du00(r) + du00(r'))/2 + (V(r) + V(r'))/2 + off-diagonal dukj terms."""
    from scipy.interpolate import CubicSpline
    if table.dy is None:
        raise ValueError("table has no beta-derivative columns")
    x = table.grid_points()
    y, dy = table.y[level_index], table.dy[level_index]
    bc = ((2, 0.0), (1, 0.0))
    q, z, s2 = (np.asarray(v, dtype=np.float64) for v in (q, z, s2))
    rmin, end = x[0], x[-1]
    r = np.clip(q + 0.5 * z, rmin, end)
    rp = np.clip(q - 0.5 * z, rmin, end)
    v0 = CubicSpline(x, y[:, 0], bc_type=bc)
    d1 = CubicSpline(x, dy[:, 1], bc_type=bc)
    out = 0.5 * (d1(r) + d1(rp)) + 0.5 * (v0(r) + v0(rp))
    use = (s2 > 0) & (q < end)
    z2 = z * z
    qc = np.minimum(q, end)
    for k in range(1, table.num_terms + 1):
        for j in range(0, k + 1):
            c = k * (k + 1) // 2 + j + 1
            out = out + np.where(use, CubicSpline(x, dy[:, c], bc_type=bc)(qc) * z2 ** j * s2 ** (k - j), 0.0)
    return out


def coulomb_bc_from_david(table: DavidTable, tau: float, num_betas: int, z1z2: float,
                          q_grid=(GRID_LOG, 0.0025, 9.5), nq: int = 200,
                          t_grid=(GRID_LINEAR, 0.0, 1.0), nt: int = 20) -> CoulombBCTable:
    """A z-independent (q, t) table from a squarer DavidPA table: U(q, t) = U_David(q, 0, (2 q t)^2)
    at tau_l = tau 2^l, l < num_betas."""
    q = _grid(*q_grid, nq)
    t = _grid(*t_grid, nt)
    taus = table.smallest_tau * 2.0 ** np.arange(table.num_tau)
    umat = np.zeros((num_betas, nq, nt))
    dumat = np.zeros((num_betas, nq, nt)) if table.dy is not None else None
    Q, T = np.meshgrid(q, t, indexing="ij")
    for b in range(num_betas):
        lev = int(np.argmin(np.abs(taus - tau * 2 ** b)))
        if abs(taus[lev] - tau * 2 ** b) > 1e-4:
            raise ValueError(f"table has no tau level {tau * 2 ** b}")
        args = (Q.ravel(), np.zeros(Q.size), (2 * Q * T).ravel() ** 2)
        umat[b] = david_u(table, lev, *args).reshape(nq, nt)
        if dumat is not None:
            dumat[b] = david_du(table, lev, *args).reshape(nq, nt)
    return CoulombBCTable(q_grid, t_grid, umat, tau, z1z2, dumat)


def coulomb_chord_table(z1z2: float, tau: float, num_betas: int, q_grid=(GRID_LOG, 0.01, 10.0), nq: int = 200,
                        t_grid=(GRID_LINEAR, 0.0, 0.95), nt: int = 20) -> CoulombBCTable:
    """Straight-line (semiclassical) action of Z1Z2/r in the (q, t) layout: with r = r' = q and
    s = 2 q t the chord distance is q sqrt(1 - 4 u(1-u) t^2), whose inverse averages to
    atanh(t) / (q t); U = tau_l Z1Z2 atanh(t) / (q t)."""
    q = _grid(*q_grid, nq)
    t = _grid(*t_grid, nt)
    g = np.where(t > 0, np.arctanh(np.minimum(t, 0.999999)) / np.where(t > 0, t, 1.0), 1.0)
    umat = np.zeros((num_betas, nq, nt))
    for b in range(num_betas):
        umat[b] = tau * 2 ** b * z1z2 * g[None, :] / q[:, None]
    dumat = np.repeat((z1z2 * g[None, :] / q[:, None])[None], num_betas, axis=0)   # linear in tau
    return CoulombBCTable(q_grid, t_grid, umat, tau, z1z2, dumat)


# --------------------------------------------------------------------------
# HDF5 coulombBCfit files (PIMC++'s IO layer; see pimcplusplus_amd.h5)
# --------------------------------------------------------------------------
def _read_grid_h5(g, path: str) -> tuple[tuple, int]:
    """((grid_type, start, end), num_points) of a grid section (ReadGrid, Common/Splines/Grid.h:654-680):
    Linear reads Start, End, NumPoints (:107-112); Log reads r0, Spacing, NumPoints and puts its
    points at r0 Spacing^i (:553-562), so its end is r0 Spacing^(n-1)."""
    gtype = g.read("Type")
    if gtype == "Linear":
        n = int(g.read("NumPoints"))
        return (GRID_LINEAR, float(g.read("Start")), float(g.read("End"))), n
    if gtype == "Log":
        n = int(g.read("NumPoints"))
        r0, spacing = float(g.read("r0")), float(g.read("Spacing"))
        return (GRID_LOG, r0, r0 * spacing ** (n - 1.0)), n
    raise ValueError(f"{path}: grid type {gtype!r} is not supported (the tables take Linear and Log grids)")


def read_coulomb_bc_h5(path: str, smallest_beta: float, num_betas: int) -> CoulombBCTable:
    """Reads a ``coulombBCfit`` pair-action file as ReadPAFit and PAcoulombBCFitClass::Read do
    (PAFit.h:30-69, PAcoulombBCFit.cc:242-311).

    Fits/Type must be coulombBCfit; Particle1 and Particle2 carry Name, lambda, Charge and Ndim;
    Potential is a Coulomb potential (Type, Z1Z2); Fits holds the qGrid and tGrid sections
    and one Fit section per beta (beta, Umat [q][t], dUmat).  Level l takes the first Fit whose
    beta is within 1e-12 (relative) of smallest_beta 2^l; a missing beta is an error, as in
    the reference (:280-285).  A Log grid keeps its points to rounding: the tables recompute
    the spacing from the end points.
    """
    from . import h5
    with h5.open_file(path) as f:
        with f.group("Fits") as fits:
            ftype = fits.read("Type")
        if ftype != "coulombBCfit":
            raise ValueError(f"{path}: pair action fit type {ftype!r}, expected 'coulombBCfit'")
        for name in ("Particle1", "Particle2"):
            with f.group(name) as pg:
                for var in ("Name", "lambda", "Charge", "Ndim"):
                    pg.read(var)
        with f.group("Potential") as pot:
            ptype = pot.read("Type")
            if ptype != "Coulomb":
                raise ValueError(f"{path}: potential type {ptype!r}, expected 'Coulomb'")
            z1z2 = float(pot.read("Z1Z2"))
        with f.group("Fits") as fits:
            with fits.group("qGrid") as g:
                q_grid, nq = _read_grid_h5(g, path)
            with fits.group("tGrid") as g:
                t_grid, nt = _read_grid_h5(g, path)
            sections = fits.sections("Fit")
            try:
                betas = [float(g.read("beta")) for _, g in sections]
                umat = np.zeros((num_betas, nq, nt))
                dumat = np.zeros((num_betas, nq, nt))
                desired = smallest_beta
                for b in range(num_betas):
                    pick = next((i for i, beta in enumerate(betas) if abs(beta - desired) / desired < 1.0e-12), None)
                    if pick is None:
                        raise ValueError(f"{path}: no Fit with beta = {desired!r} (have {betas})")
                    g = sections[pick][1]
                    for out, var in ((umat, "Umat"), (dumat, "dUmat")):
                        m = np.asarray(g.read(var), dtype=np.float64)
                        if m.shape != (nq, nt):
                            raise ValueError(f"{path}: Fit {var} has shape {m.shape}, expected {(nq, nt)}")
                        out[b] = m
                    desired *= 2.0
            finally:
                for _, g in sections:
                    g.close()
    return CoulombBCTable(q_grid, t_grid, umat, smallest_beta, z1z2, dumat)
