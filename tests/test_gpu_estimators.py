"""GPU estimators against the oracle (FP64, 1e-12; histograms exact), and the
thermodynamic energy of free particles against its closed form 3N/(2 beta)."""

import math

import numpy as np
import pytest

from pimcplusplus_amd import abi, tables, workloads
from tests.helpers import GOLDEN
from tests.test_oracle_estimators import free_table

pytestmark = pytest.mark.gpu


def both(oracle, gpu, system):
    return workloads.load(oracle, system), workloads.load(gpu, system)


def close(a, b, rel=1e-12):
    return abs(a - b) <= rel * (1.0 + abs(b))


def check_energy(oracle, gpu, system, moves=30):
    io, ig = both(oracle, gpu, system)
    oracle.bisection_run(io, system.num_levels, system.seed, 0, moves)
    gpu.bisection_run(ig, system.num_levels, system.seed, 0, moves, abi.FP64)
    for mode in (abi.OLDMODE, abi.NEWMODE):
        eo = oracle.energy(io, mode)
        eg = gpu.energy(ig, mode)
        for a, b in zip(eg, eo):
            for f in ("kinetic", "du_short", "du_long", "total"):
                assert close(getattr(a, f), getattr(b, f)), (f, getattr(a, f), getattr(b, f))
            for f in ("v_short", "v_long"):
                if math.isnan(getattr(b, f)):
                    assert math.isnan(getattr(a, f)), f
                else:
                    assert close(getattr(a, f), getattr(b, f)), (f, getattr(a, f), getattr(b, f))
    return eo


def test_energy_lj(oracle, gpu):
    check_energy(oracle, gpu, workloads.config0(num_replicas=3))


def test_energy_charged_lj_long_range(oracle, gpu):
    e = check_energy(oracle, gpu, workloads.charged_lj_system(n_plus=12, n_minus=12, m=16, num_replicas=2))
    assert e[0].du_long != 0.0


def test_energy_hydrogen(oracle, gpu):
    t = tables.load_npz(f"{GOLDEN}/testcoul_fit2.npz")
    e = check_energy(oracle, gpu, workloads.hydrogen_system(t, num_replicas=2, n_pairs=8, m=32, num_levels=3))
    assert math.isfinite(e[0].v_short) and e[0].v_long != 0.0 and e[0].du_long != 0.0


def test_energy_hydrogen_ewald_leaves_the_potential_unknown(oracle, gpu):
    t = tables.load_npz(f"{GOLDEN}/testcoul_fit2.npz")
    e = check_energy(oracle, gpu, workloads.hydrogen_system(t, num_replicas=1, n_pairs=4, m=16, num_levels=2,
                                                            long_range="ewald"), moves=10)
    assert math.isnan(e[0].v_short) and math.isnan(e[0].v_long)


def test_energy_ocp_background(oracle, gpu):
    e = check_energy(oracle, gpu, workloads.ocp_system(num_replicas=2, n=64, m=16, num_levels=3, kc_units=4.0))
    assert math.isfinite(e[0].v_short) and e[0].v_long != 0.0


# 101 points: every block counts in its LDS histogram; 12001 points (12000 bins): too many bins
# for LDS, each count goes to the global histogram.
@pytest.mark.parametrize("points", [101, 12001])
def test_pair_correlation_matches_oracle(oracle, gpu, points):
    sysm = workloads.charged_lj_system(n_plus=12, n_minus=20, m=16, num_replicas=1)
    io, ig = both(oracle, gpu, sysm)
    oracle.bisection_run(io, sysm.num_levels, sysm.seed, 0, 40, species=1)
    gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, 40, abi.FP64, species=1)
    L = sysm.replicas[0].box[0]
    for s1, s2 in ((0, 0), (1, 1), (0, 1)):
        ho = oracle.pair_correlation(io[0], s1, s2, 0.0, L / 2, points)
        hg = gpu.pair_correlation(ig[0], s1, s2, 0.0, L / 2, points)
        assert np.array_equal(ho, hg) and ho.sum() > 0


def test_pair_correlation_of_the_large_ocp_matches_oracle(oracle, gpu):
    """N = 1024, M = 128: 67 M pair distances over the LDS histograms of 4096 blocks."""
    sysm = workloads.ocp_system(num_replicas=1)
    io, ig = both(oracle, gpu, sysm)
    L = sysm.replicas[0].box[0]
    ho = oracle.pair_correlation(io[0], 0, 0, 0.0, L / 2, 41)
    hg = gpu.pair_correlation(ig[0], 0, 0, 0.0, L / 2, 41)
    assert np.array_equal(ho, hg) and ho.sum() > 0


def test_free_particle_energy_is_three_n_over_two_beta(gpu):
    # distinguishable free particles, thermal wavelength << box: <E> = 3N/(2 beta)
    n, m, tau, lam, box, nrep = 6, 16, 0.1, 0.5, 40.0, 512
    t = free_table(box, smallest_tau=tau, num_tau=4)
    rng = np.random.default_rng(7)
    reps = []
    for r in range(nrep):
        beads = np.repeat(rng.uniform(-box / 2, box / 2, size=(1, n, 3)), m + 1, axis=0)
        reps.append(workloads.ReplicaSpec(num_particles=n, num_time_slices=m, box=(box, box, box), tau=tau,
                                          species_count=[n], lambdas=[lam], fit_keys=["free"], beads=beads,
                                          rng_stream=r))
    sysm = workloads.System("free", {"free": t}, {"free": tau}, reps, 3, 4242)
    ids = workloads.load(gpu, sysm)
    move, samples = 0, []
    for block in range(60):
        gpu.bisection_run(ids, 3, sysm.seed, move, 40, abi.FP64, records=False, wait=False)
        move += 40
        if block >= 20:
            samples.append([e.total / m for e in gpu.energy(ids)])
    s = np.array(samples)               # [samples][replicas]
    per_rep = s.mean(axis=0)            # replicas are independent
    mean, err = per_rep.mean(), per_rep.std(ddof=1) / math.sqrt(nrep)
    exact = 3.0 * n / (2.0 * m * tau)
    assert abs(mean - exact) < 4.0 * err, (mean, err, exact)
    assert err < 0.05 * exact


def test_mixed_and_fp64_energies_agree_within_error_bars(gpu):
    """The thermodynamic energy and g(r) of mixed and FP64 chains that share no random numbers
    (the mixed chains run on another seed): each replica starts from the same configuration in
    both precisions, so the replicas' (mixed - FP64) differences are compared with their own
    standard error -- the energy within 4 of it, every g(r) bin within 4.5."""
    sysm = workloads.config0(num_replicas=256)
    L = sysm.replicas[0].box[0]
    bins, blocks, moves, drop = 30, 8, 100, 2
    res = {}
    for prec, seed in ((abi.FP64, sysm.seed), (abi.MIXED, sysm.seed + 7919)):
        ids = workloads.load(gpu, sysm)
        e, g = [], []
        for block in range(blocks):
            gpu.bisection_run(ids, sysm.num_levels, seed, moves * block, moves, prec, records=False, wait=False)
            e.append([x.total for x in gpu.energy(ids)])
            g.append([gpu.pair_correlation(r, 0, 0, 0.0, L / 2, bins + 1) for r in ids])
        res[prec] = (np.array(e[drop:]).mean(axis=0), np.array(g[drop:], dtype=float).mean(axis=0))
    (ea, ga), (eb, gb) = res[abi.FP64], res[abi.MIXED]
    assert not np.array_equal(ea, eb)          # the chains differ
    de, dg = eb - ea, gb - ga
    n = len(de)
    assert abs(de.mean()) < 4.0 * de.std(ddof=1) / math.sqrt(n)
    gerr = dg.std(axis=0, ddof=1) / math.sqrt(n)
    ok = gerr > 0
    assert ok.sum() >= bins // 2
    assert np.max(np.abs(dg.mean(axis=0)[ok]) / gerr[ok]) < 4.5


@pytest.mark.parametrize("name", ["lj", "charged_lj", "hydrogen"])
def test_d_dbeta_per_range_matches_oracle(oracle, gpu, name):
    from tests.test_oracle_dbeta import LR, systems
    system = dict(systems())[name]
    io, ig = both(oracle, gpu, system)
    oracle.bisection_run(io, system.num_levels, system.seed, 0, 25)
    gpu.bisection_run(ig, system.num_levels, system.seed, 0, 25, abi.FP64)
    M = system.replicas[0].num_time_slices
    rng = np.random.default_rng(len(name))
    reqs = []
    for _ in range(24):
        level = int(rng.integers(0, 3))
        skip = 1 << level
        n = int(rng.integers(1, M // skip + 1))
        s1 = int(rng.integers(0, M - n * skip + 1))
        reqs.append((int(rng.integers(0, 2)), int(rng.integers(0, 2)), s1, s1 + n * skip, level))
    for action in (abi.ACTION_SHORT_RANGE, abi.ACTION_KINETIC, abi.ACTION_LONG_RANGE):
        if action == LR and name == "lj":
            continue
        print("d_dBeta", name, action, flush=True)
        want = oracle.d_dbeta(action, [(io[x[0]],) + x[1:] for x in reqs])
        got = gpu.d_dbeta(action, [(ig[x[0]],) + x[1:] for x in reqs])
        assert np.max(np.abs(got - want) / (1 + np.abs(want))) < 1e-12, action
