"""Synthetic PIMC systems of the benchmark configurations.

Each builder returns plain data (tables, replica parameters, initial beads) so
the same system can be loaded into any library exporting the pimc_* ABI.

Units: Kelvin and Angstrom for the Lennard-Jones systems.  lambda = hbar^2/2m
for a 4 amu particle is 24.2542/4 K A^2.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

import math

from .breakup import optimized_breakup, quintic_data
from .ewald import dulong_quintic, ewald_dbeta_tables, ewald_tables, ulong_quintic
from .tables import (GRID_LINEAR, GRID_LOG, CoulombBCTable, DavidTable, coulomb_bc_from_david,
                     coulomb_chord_table, lj_david_table)

LAMBDA_AMU = 24.2542          # hbar^2 / (2 * 1 amu * k_B) in K A^2
LJ_EPS = 10.22                # K
LJ_SIGMA = 2.556              # A
LJ_DENSITY = 0.02186          # A^-3


@dataclass
class ReplicaSpec:
    num_particles: int
    num_time_slices: int
    box: tuple
    tau: float
    species_count: list
    lambdas: list
    fit_keys: list             # [ns*ns] keys into System.tables
    beads: np.ndarray          # [(M+1)][N][3]
    rng_stream: int = 0
    num_images: int = 0
    k_cutoff: float = 0.0
    use_long_range: bool = False
    use_background: bool = False
    # (sa, sb) -> (z1z2, alpha): closed-form Ewald long-range tables on the replica's k set
    ewald: dict = field(default_factory=dict)
    # (sa, sb) -> (z1z2, alpha): Ulong(r) = tau_l z1z2 erf(alpha r)/r subtracted from long-ranged
    # pair fits, a quintic on LinearGrid(0, 0.75 |box|, 1000)
    ulong: dict = field(default_factory=dict)
    # (sa, sb) -> key of a CoulombBCTable in System.tables: the optimized breakup of that
    # pair's action gives both the k-space tables and Ulong(r)
    breakup: dict = field(default_factory=dict)


@dataclass
class System:
    name: str
    tables: dict               # key -> DavidTable or CoulombBCTable
    desired_tau: dict          # key -> tau the fit is selected at (DavidTable keys)
    replicas: list             # [ReplicaSpec]
    num_levels: int
    seed: int
    species: int = 0
    meta: dict = field(default_factory=dict)


def lj_box(n: int, density: float = LJ_DENSITY) -> float:
    return (n / density) ** (1.0 / 3.0)


def random_start(n: int, box: float, nslices: int, rng: np.random.Generator,
                 jitter: float = 0.1) -> np.ndarray:
    """A classical start, identical on every slice: n sites drawn at random from the
    smallest fcc lattice (4k^3 sites) that holds them, each jittered uniformly by
    +-jitter, in [-L/2, L/2)^3.  A uniform random start at this density places pairs
    deep in the repulsive core; the lattice is the low-energy equivalent of the
    reference's RANDOMFIXED initial paths (PIMC++/src/InitPaths.cc:967-999)."""
    k = 1
    while 4 * k ** 3 < n:
        k += 1
    a = box / k
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3)
    sites = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) + 0.25) * a - 0.5 * box
    pick = rng.permutation(sites.shape[0])[:n]
    pos = sites[pick] + rng.uniform(-jitter, jitter, size=(n, 3))
    pos -= box * np.floor(pos / box + 0.5)
    return np.repeat(pos[None, :, :], nslices + 1, axis=0).copy()


def lj_table_for(box: float, tau: float, num_levels: int, num_points: int = 200,
                 r_min: float = 0.6 * LJ_SIGMA) -> DavidTable:
    return lj_david_table(LJ_EPS, LJ_SIGMA, r_min, box / 2.0, num_points, tau, num_levels)


def lj_system(name: str, n: int, m: int, tau: float, num_levels: int, seed: int,
              num_replicas: int = 1) -> System:
    """num_replicas independent copies of one LJ system (different starts and random streams)."""
    box = lj_box(n)
    table = lj_table_for(box, tau, num_levels)
    rng = np.random.default_rng(seed)
    lam = LAMBDA_AMU / 4.0
    reps = []
    for r in range(num_replicas):
        reps.append(ReplicaSpec(num_particles=n, num_time_slices=m, box=(box, box, box), tau=tau,
                                species_count=[n], lambdas=[lam], fit_keys=["lj"],
                                beads=random_start(n, box, m, rng), rng_stream=r))
    return System(name=name, tables={"lj": table}, desired_tau={"lj": tau}, replicas=reps,
                  num_levels=num_levels, seed=seed)


def dilute_boson_system(num_replicas: int = 2, n: int = 16, m: int = 32, tau: float = 0.05, num_levels: int = 3,
                        density: float = 0.005, seed: int = 77) -> System:
    """Helium-4 atoms (mass 4 amu, LJ pair action) at a low density, for bisection moves with
    PermuteType TABLE: nearest neighbours sit ~6 A apart, so over a window of 2^L slices
    (4 lambda beta ~ 10 A^2 at the defaults) exchanges are drawn and accepted regularly."""
    box = (n / density) ** (1.0 / 3.0)
    table = lj_table_for(box, tau, num_levels)
    rng = np.random.default_rng(seed)
    reps = [ReplicaSpec(num_particles=n, num_time_slices=m, box=(box, box, box), tau=tau, species_count=[n],
                        lambdas=[LAMBDA_AMU / 4.0], fit_keys=["lj"], beads=random_start(n, box, m, rng),
                        rng_stream=r)
            for r in range(num_replicas)]
    return System(name=f"bose-n{n}-m{m}-l{num_levels}", tables={"lj": table}, desired_tau={"lj": tau},
                  replicas=reps, num_levels=num_levels, seed=seed)


def coulomb_boson_system(num_replicas: int = 2, n: int = 12, m: int = 32, tau: float = 0.125, num_levels: int = 3,
                         density: float = 0.05, z1z2: float = 0.5, seed: int = 5) -> System:
    """Bosons of mass 1 (lambda = 1/2 in Hartree units) with a weak repulsive Coulomb pair
    action (a PAcoulombBCFitClass table on a log q grid) and no long-range action: the generic
    chain kernel with PermuteType TABLE, where 2-, 3- and 4-cycles are all accepted."""
    from .tables import GRID_LOG, coulomb_chord_table
    box = (n / density) ** (1.0 / 3.0)
    tab = coulomb_chord_table(z1z2, tau, num_levels, q_grid=(GRID_LOG, 0.005, 0.75 * math.sqrt(3.0) * box))
    rng = np.random.default_rng(seed)
    reps = [ReplicaSpec(num_particles=n, num_time_slices=m, box=(box, box, box), tau=tau, species_count=[n],
                        lambdas=[0.5], fit_keys=["bb"], beads=random_start(n, box, m, rng), rng_stream=r)
            for r in range(num_replicas)]
    return System(name=f"coulomb-bosons-n{n}-m{m}-l{num_levels}", tables={"bb": tab}, desired_tau={},
                  replicas=reps, num_levels=num_levels, seed=seed)


def config0(num_replicas: int = 1) -> System:
    """64 LJ particles, M=32, tau=0.025, level-3 bisection, seed 1000."""
    return lj_system("lj-n64-m32-l3", 64, 32, 0.025, 3, 1000, num_replicas)


def config1(num_replicas: int = 1) -> System:
    """256 LJ particles, M=64, tau=0.0125, level-4 bisection, seed 1001."""
    return lj_system("lj-n256-m64-l4", 256, 64, 0.0125, 4, 1001, num_replicas)


def config3(num_replicas: int = 128, seed: int = 1003) -> System:
    """Uneven LJ replicas: N log-uniform in [32, 512], M uniform in {16, 32, 64, 128},
    tau=0.025, level-3 bisection.  Each replica has its own box (density fixed) and table."""
    rng = np.random.default_rng(seed)
    tau, num_levels = 0.025, 3
    lam = LAMBDA_AMU / 4.0
    tables, dtau, reps = {}, {}, []
    for r in range(num_replicas):
        n = int(round(np.exp(rng.uniform(np.log(32), np.log(512)))))
        m = int(rng.choice([16, 32, 64, 128]))
        box = lj_box(n)
        key = f"lj{n}"
        if key not in tables:
            tables[key] = lj_table_for(box, tau, num_levels)
            dtau[key] = tau
        reps.append(ReplicaSpec(num_particles=n, num_time_slices=m, box=(box, box, box), tau=tau,
                                species_count=[n], lambdas=[lam], fit_keys=[key],
                                beads=random_start(n, box, m, rng), rng_stream=r))
    return System(name="lj-uneven", tables=tables, desired_tau=dtau, replicas=reps,
                  num_levels=num_levels, seed=seed)


def load(lib, system: System):
    """Creates the system's fits and replicas in `lib` (a PimcLib).  Returns replica ids."""
    fit_ids = {}
    for key, t in system.tables.items():
        if isinstance(t, CoulombBCTable):
            fit_ids[key] = lib.fit_create_coulomb_bc(t.q_grid, t.t_grid, t.umat, t.smallest_beta, t.z1z2,
                                                     dumat=t.dumat)
        else:
            fit_ids[key] = lib.fit_create_david(t.grid_type, t.grid_start, t.grid_end, t.y, t.num_terms,
                                                t.smallest_tau, system.desired_tau[key], dy=t.dy)
    ids = []
    for spec in system.replicas:
        rid = lib.replica_create(spec.num_particles, spec.num_time_slices, spec.box, spec.tau,
                                 spec.species_count, spec.lambdas,
                                 [fit_ids[k] for k in spec.fit_keys], num_images=spec.num_images,
                                 k_cutoff=spec.k_cutoff, use_long_range=spec.use_long_range,
                                 use_background=spec.use_background, rng_stream=spec.rng_stream)
        lib.path_set(rid, spec.beads)
        if spec.ulong:
            r_end = 0.75 * math.sqrt(sum(b * b for b in spec.box))
            for (sa, sb), (z1z2, alpha) in spec.ulong.items():
                start, end, coefs = ulong_quintic(z1z2, alpha, spec.tau, system.num_levels, r_end)
                _, _, dcoefs = dulong_quintic(z1z2, alpha, spec.tau, system.num_levels, r_end)
                lib.set_ulong(rid, sa, sb, start, end, coefs, dcoefs)
        if spec.breakup:
            kv, _ = lib.kvecs(rid)
            for (sa, sb), key in spec.breakup.items():
                tab = system.tables[key]
                u, du, v, qu, qdu, qv = _breakup_tables(tab, spec.box, spec.k_cutoff, kv, system.num_levels)
                lib.set_ulong(rid, sa, sb, 0.0, float(u.r_grid[-1]), qu, qdu)
                lib.set_long_range(rid, sa, sb, u.ulong_k, u.ulong_r0, u.ushort_k0, u.ulong_k0,
                                   du.ulong_k, du.ulong_r0, du.ushort_k0, du.ulong_k0)
                lib.set_vlong(rid, sa, sb, v.ulong_k[0], v.ulong_r0[0], v.ushort_k0[0], v.ulong_k0[0],
                              0.0, float(v.r_grid[-1]), qv[0])
            lib.rho_k_update_all(rid)
        if spec.ewald:
            kv, _ = lib.kvecs(rid)
            for (sa, sb), (z1z2, alpha) in spec.ewald.items():
                lib.set_long_range(rid, sa, sb,
                                   *ewald_tables(kv, spec.box, z1z2, alpha, spec.tau, system.num_levels),
                                   *ewald_dbeta_tables(kv, spec.box, z1z2, alpha, spec.tau, system.num_levels))
            lib.rho_k_update_all(rid)
        ids.append(rid)
    return ids


def charged_lj_system(n_plus: int = 24, n_minus: int = 24, m: int = 32, tau: float = 0.025,
                      num_levels: int = 3, kc_units: float = 4.0, seed: int = 1100,
                      num_replicas: int = 1, charge_scale: float = 2.0) -> System:
    """Two LJ species carrying charges +-q with an Ewald long-range action (test system for the
    k-space path): short range is the LJ pair action, long range the Ewald k sum of q^2/r in
    K*A units (q^2 = charge_scale K A)."""
    n = n_plus + n_minus
    box = lj_box(n)
    table = lj_table_for(box, tau, num_levels)
    rng = np.random.default_rng(seed)
    lam = LAMBDA_AMU / 4.0
    kc = kc_units * 2.0 * np.pi / box
    alpha = 5.0 / box
    reps = []
    for r in range(num_replicas):
        reps.append(ReplicaSpec(num_particles=n, num_time_slices=m, box=(box, box, box), tau=tau,
                                species_count=[n_plus, n_minus], lambdas=[lam, lam],
                                fit_keys=["lj"] * 4, beads=random_start(n, box, m, rng), rng_stream=r,
                                k_cutoff=kc, use_long_range=True,
                                ewald={(0, 0): (charge_scale, alpha), (1, 1): (charge_scale, alpha),
                                       (0, 1): (-charge_scale, alpha)}))
    return System(name="charged-lj", tables={"lj": table}, desired_tau={"lj": tau}, replicas=reps,
                  num_levels=num_levels, seed=seed)


# --------------------------------------------------------------------------
# Coulomb systems (Hartree atomic units)
# --------------------------------------------------------------------------
LAMBDA_E = 0.5                 # hbar^2 / 2 m_e
PROTON_MASS = 1836.15267


def coulomb_box(n: int, rs: float) -> float:
    return (4.0 * math.pi * n / 3.0) ** (1.0 / 3.0) * rs


def kcut(units: float, box: float) -> float:
    """k cutoff units * 2 pi / L, placed just inside the shell |n|^2 = units^2 so the
    strict |k|^2 < kc^2 test of the k set is not decided by rounding."""
    return (units - 1e-9) * 2.0 * math.pi / box


_BREAKUP_CACHE: dict = {}


def _breakup_tables(tab, box, kc: float, kvecs: np.ndarray, num_levels: int):
    """Optimized breakups of U, dU/dbeta and V of a Coulomb pair table for one box and k set
    (LongRangeClass::OptimizedBreakup_U / _dU / _V) and the quintic data of their real-space
    parts, cached across replicas with the same box."""
    key = (id(tab), tuple(box), kc, num_levels, kvecs.shape[0])
    if key not in _BREAKUP_CACHE:
        q_end = float(tab.grid_points("q")[-1])
        z1z2 = float(tab.z1z2)

        def coulomb(r):
            return z1z2 / r

        u = optimized_breakup([tab.udiag(l) for l in range(num_levels)], box, kc, kvecs, tail_end=q_end)
        du = optimized_breakup([tab.udiag(l, derivative=True) for l in range(num_levels)], box, kc, kvecs,
                               tail_end=q_end, short_of=coulomb)
        v = optimized_breakup([coulomb], box, kc, kvecs, tail_coef=(z1z2, 0.0, 0.0))
        _BREAKUP_CACHE[key] = (u, du, v, quintic_data(u.r_grid, u.ulong_r), quintic_data(du.r_grid, du.ulong_r),
                               quintic_data(v.r_grid, v.ulong_r))
    return _BREAKUP_CACHE[key]


def _long_range_tables(kind: str, z: list, alpha: float, keys: dict) -> dict:
    if kind == "breakup":
        return {"breakup": keys}
    if kind == "ewald":
        return {"ewald": _coulomb_pairs(z, alpha), "ulong": _coulomb_pairs(z, alpha)}
    raise ValueError(f"long_range must be 'breakup' or 'ewald', not {kind!r}")


def _coulomb_pairs(z: list, alpha: float) -> dict:
    ns = len(z)
    return {(a, b): (z[a] * z[b], alpha) for a in range(ns) for b in range(a, ns)}


def hydrogen_system(dm_table: DavidTable, num_replicas: int = 1, n_pairs: int = 32, rs: float = 1.75,
                    m: int = 128, tau: float = 0.125, num_levels: int = 4, kc_units: float = 5.0,
                    seed: int = 1002, long_range: str = "breakup") -> System:
    """Hydrogen plasma: n_pairs electrons (species 0, moved) and n_pairs protons (species 1).
    e-p: the squarer table `dm_table` (testcoul.dm, Z1Z2 = -1) resampled on the
    PAcoulombBCFitClass (q, t) layout; e-e and p-p: straight-line Coulomb tables.  Every pair
    is long-ranged: its long part Ulong is subtracted in the short-range action and the
    k-space action is added at the lowest bisection level.  long_range "breakup": the
    optimized breakup of each pair table (as PIMC++); "ewald": the closed-form Ewald split
    erf(alpha r)/r with alpha = kc / 6."""
    n = 2 * n_pairs
    box = coulomb_box(n_pairs, rs)
    kc = kcut(kc_units, box)
    alpha = kc / 6.0
    q_end = 0.75 * math.sqrt(3.0) * box
    tabs = {
        "ee": coulomb_chord_table(1.0, tau, num_levels, q_grid=(GRID_LOG, 0.005, q_end)),
        "pp": coulomb_chord_table(1.0, tau, num_levels, q_grid=(GRID_LOG, 0.005, q_end)),
        "ep": coulomb_bc_from_david(dm_table, tau, num_levels, -1.0,
                                    q_grid=(GRID_LOG, 0.0025, min(q_end, dm_table.grid_end * 0.95))),
    }
    rng = np.random.default_rng(seed)
    reps = []
    for r in range(num_replicas):
        beads = random_start(n, box, m, rng, jitter=0.05 * box)
        reps.append(ReplicaSpec(num_particles=n, num_time_slices=m, box=(box, box, box), tau=tau,
                                species_count=[n_pairs, n_pairs],
                                lambdas=[LAMBDA_E, LAMBDA_E / PROTON_MASS],
                                fit_keys=["ee", "ep", "ep", "pp"], beads=beads, rng_stream=r,
                                k_cutoff=kc, use_long_range=True, **_long_range_tables(
                                    long_range, [-1.0, 1.0], alpha, {(0, 0): "ee", (0, 1): "ep", (1, 1): "pp"})))
    return System(name=f"hydrogen-n{n}-m{m}-l{num_levels}", tables=tabs, desired_tau={}, replicas=reps,
                  num_levels=num_levels, seed=seed,
                  meta={"rs": rs, "box": box, "kc": kc, "alpha": alpha, "tau": tau})


def config2(dm_table: DavidTable, num_replicas: int = 1, m: int = 128) -> System:
    """32 e + 32 p, rs=1.75 (L = 8.9561 bohr), tau=0.125 Ha^-1, kc = 5 2pi/L (242 k).

    The e-p table tabulates tau = 2^i / 64 Ha^-1 only, so tau stays 0.125 and the temperature
    follows from M: T = 1 / (M tau) Ha is 19,737 K at M = 128 and 15,037 K at M = 168."""
    return hydrogen_system(dm_table, num_replicas=num_replicas, m=m)


HARTREE_K = 315775.02   # 1 Ha / k_B in kelvin


def temperature_k(system: System) -> float:
    """T = 1 / (M tau) of the first replica in kelvin: tau is in Ha^-1 for the Coulomb systems
    (those with an rs) and in K^-1 for the Lennard-Jones ones."""
    r = system.replicas[0]
    return (HARTREE_K if "rs" in system.meta else 1.0) / (r.num_time_slices * r.tau)


def ocp_system(num_replicas: int = 16, n: int = 1024, rs: float = 1.0, m: int = 128, tau: float = 0.05,
               num_levels: int = 4, kc_units: float = 10.0, seed: int = 2000, long_range: str = "breakup") -> System:
    """One-component plasma with a neutralising background: n electrons, straight-line Coulomb
    pair table, Ulong subtraction, k-space action with the background terms (long_range as in
    hydrogen_system)."""
    box = coulomb_box(n, rs)
    kc = kcut(kc_units, box)
    alpha = kc / 6.0
    q_end = 0.75 * math.sqrt(3.0) * box
    tabs = {"ee": coulomb_chord_table(1.0, tau, num_levels, q_grid=(GRID_LOG, 0.005, q_end))}
    reps = []
    for r in range(num_replicas):
        rng = np.random.default_rng(seed + r)
        reps.append(ReplicaSpec(num_particles=n, num_time_slices=m, box=(box, box, box), tau=tau,
                                species_count=[n], lambdas=[LAMBDA_E], fit_keys=["ee"],
                                beads=random_start(n, box, m, rng, jitter=0.05 * box), rng_stream=r,
                                k_cutoff=kc, use_long_range=True, use_background=True,
                                **_long_range_tables(long_range, [1.0], alpha, {(0, 0): "ee"})))
    return System(name=f"ocp-n{n}-m{m}-l{num_levels}", tables=tabs, desired_tau={}, replicas=reps,
                  num_levels=num_levels, seed=seed,
                  meta={"rs": rs, "box": box, "kc": kc, "alpha": alpha, "tau": tau})


def config4(num_replicas: int = 16) -> System:
    """OCP N=1024, rs=1 (L = 16.2479 bohr), M=128, kc = 10 2pi/L (2069 k), 16 replicas."""
    return ocp_system(num_replicas=num_replicas)
