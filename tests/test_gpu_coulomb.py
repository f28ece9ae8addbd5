"""GPU parity of the Coulomb path: PAcoulombBCFitClass::U (bicubic), the Ulong subtraction of
ShortRangeClass::SingleAction, and bisection chains of the hydrogen plasma and the
one-component plasma (short range + Ewald long range) against the oracle.

FP64 path: 1e-12 relative.  Mixed path (FP32 bicubic cells away from the core, FP32
Horner tail of the quintic Ulong): the contract is |dS| <= 1e-6 per stage; single
evaluations are checked to 2e-6 (1 + |U|) and must differ from FP64 somewhere (the FP32
path ran)."""

import math

import numpy as np
import pytest

from pimcplusplus_amd import abi, ewald, tables, workloads
from tests.helpers import GOLDEN
from tests.test_gpu_parity import compare_chains
from tests.test_oracle_coulomb import smooth_umat

pytestmark = pytest.mark.gpu

FP64, MIXED = abi.FP64, abi.MIXED


def _testcoul():
    return tables.load_npz(f"{GOLDEN}/testcoul_fit2.npz")


@pytest.mark.parametrize("qtype", [abi.GRID_LOG, abi.GRID_LINEAR])
def test_fit_eval_u_coulomb_bc(oracle, gpu, qtype):
    rng = np.random.default_rng(300 + qtype)
    qg, tg = (qtype, 0.01, 6.0), (abi.GRID_LINEAR, 0.0, 1.0)
    umat = smooth_umat(3, tables._grid(*qg, 90), tables._grid(*tg, 17))
    fo = oracle.fit_create_coulomb_bc(qg, tg, umat, 0.125, -1.0)
    fg = gpu.fit_create_coulomb_bc(qg, tg, umat, 0.125, -1.0)
    n = 20000
    q = rng.uniform(0.0, 7.0, n)
    t = rng.uniform(0.0, 1.0, n)
    q[:4] = [0.0, 6.0, 6.0000001, 0.01]
    s2 = (2 * q * t) ** 2
    z = rng.uniform(-1, 1, n)
    for lev in range(3):
        want = oracle.fit_eval_u(fo, lev, q, z, s2)
        got = gpu.fit_eval_u(fg, lev, q, z, s2, FP64)
        # contracted (FMA) Hermite sums: ~1e-13 where q=0 extrapolates below the grid
        assert np.max(np.abs(got - want) / (1 + np.abs(want))) < 1e-12
        mixed = gpu.fit_eval_u(fg, lev, q, z, s2, MIXED)
        assert np.max(np.abs(mixed - want) / (1 + np.abs(want))) < 2e-6
    # a table small enough for FP32 cells: the mixed path evaluates them in FP32
    small = 0.02 * umat
    fo = oracle.fit_create_coulomb_bc(qg, tg, small, 0.125, -1.0)
    fg = gpu.fit_create_coulomb_bc(qg, tg, small, 0.125, -1.0)
    for lev in range(3):
        want = oracle.fit_eval_u(fo, lev, q, z, s2)
        got = gpu.fit_eval_u(fg, lev, q, z, s2, FP64)
        assert np.max(np.abs(got - want) / (1 + np.abs(want))) < 1e-12
        mixed = gpu.fit_eval_u(fg, lev, q, z, s2, MIXED)
        assert np.max(np.abs(mixed - want)) < 2e-7
        assert np.mean(mixed != got) > 0.5


def test_fit_eval_u_testcoul_resampled(oracle, gpu):
    bc = tables.coulomb_bc_from_david(_testcoul(), 0.125, 3, -1.0)
    fo = oracle.fit_create_coulomb_bc(bc.q_grid, bc.t_grid, bc.umat, bc.smallest_beta, bc.z1z2)
    fg = gpu.fit_create_coulomb_bc(bc.q_grid, bc.t_grid, bc.umat, bc.smallest_beta, bc.z1z2)
    rng = np.random.default_rng(301)
    q = np.exp(rng.uniform(np.log(0.002), np.log(12.0), 20000))
    s2 = (2 * q * rng.uniform(0, 1, q.size)) ** 2
    for lev in range(3):
        want = oracle.fit_eval_u(fo, lev, q, np.zeros_like(q), s2)
        got = gpu.fit_eval_u(fg, lev, q, np.zeros_like(q), s2, FP64)
        assert np.max(np.abs(got - want) / (1 + np.abs(want))) < 1e-12


def load_both(oracle, gpu, system):
    return workloads.load(oracle, system), workloads.load(gpu, system)


def test_short_range_with_ulong_requests(oracle, gpu):
    sysm = workloads.hydrogen_system(_testcoul(), num_replicas=2, n_pairs=12, m=32, num_levels=3)
    io, ig = load_both(oracle, gpu, sysm)
    rng = np.random.default_rng(302)
    n = sysm.replicas[0].num_particles
    for r, (a, b) in enumerate(zip(io, ig)):
        new = sysm.replicas[r].beads.copy()
        new[5:20, [1, 13, 20]] += rng.normal(scale=0.2, size=(15, 3, 3))
        oracle.path_set(a, new, abi.NEWMODE)
        gpu.path_set(b, new, abi.NEWMODE)
    reqs = []
    for _ in range(80):
        level = int(rng.integers(0, 3))
        span = 8 << (level > 1)
        s1 = int(rng.integers(0, 33 - span))
        ptcls = [int(p) for p in rng.choice(n, size=int(rng.integers(1, 5)), replace=False)]
        reqs.append((int(rng.integers(0, 2)), int(rng.integers(0, 2)), s1, s1 + span, level, ptcls))
    want = oracle.action_short_range([(io[x[0]],) + x[1:] for x in reqs])
    got = gpu.action_short_range([(ig[x[0]],) + x[1:] for x in reqs], FP64)
    assert np.max(np.abs(got - want) / (1 + np.abs(want))) < 1e-12
    mixed = gpu.action_short_range([(ig[x[0]],) + x[1:] for x in reqs], MIXED)
    assert np.max(np.abs(mixed - want)) < 1e-6
    assert np.any(mixed != got)


def test_bisection_hydrogen_fp64(oracle, gpu):
    sysm = workloads.hydrogen_system(_testcoul(), num_replicas=2, n_pairs=16, m=64, num_levels=4)
    io, ig = load_both(oracle, gpu, sysm)
    ro, so = oracle.bisection_run(io, sysm.num_levels, sysm.seed, 0, 200)
    rg, sg = gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, 200, FP64)
    compare_chains(ro, rg)
    assert so.accepted == sg.accepted > 0 and so.k_evals == sg.k_evals > 0
    for a, b in zip(io, ig):
        assert np.max(np.abs(oracle.rho_k(a) - gpu.rho_k(b))) < 1e-10
        assert np.max(np.abs(oracle.path_get(a) - gpu.path_get(b))) < 1e-12


def test_bisection_hydrogen_mixed_delta_s(oracle, gpu):
    sysm = workloads.hydrogen_system(_testcoul(), num_replicas=2, n_pairs=16, m=64, num_levels=4)
    io, ig = load_both(oracle, gpu, sysm)
    ro, _ = oracle.bisection_run(io, sysm.num_levels, sysm.seed, 0, 300)
    rm, _ = gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, 300, MIXED)
    rg = workloads.load(gpu, sysm)
    r64, _ = gpu.bisection_run(rg, sysm.num_levels, sysm.seed, 0, 300, FP64)
    worst, n = 0.0, 0
    for r in range(2):
        for m in range(300):
            ns = ro["num_stages"][r, m]
            if ns != rm["num_stages"][r, m] or ro["accepted"][r, m] != rm["accepted"][r, m]:
                break
            d = (ro["new_action"][r, m, :ns] - ro["old_action"][r, m, :ns]) - \
                (rm["new_action"][r, m, :ns] - rm["old_action"][r, m, :ns])
            worst = max(worst, float(np.max(np.abs(d))))
            n += ns
    print(f"hydrogen mixed: max |dS error| {worst:.3g} over {n} stages", flush=True)
    assert n > 300 and worst <= 1e-6
    assert not np.array_equal(rm["new_action"], r64["new_action"])


def test_bisection_ocp_background_fp64_and_mixed(oracle, gpu):
    sysm = workloads.ocp_system(num_replicas=2, n=128, m=32, num_levels=3, kc_units=6.0, seed=2100)
    io, ig = load_both(oracle, gpu, sysm)
    ro, _ = oracle.bisection_run(io, sysm.num_levels, sysm.seed, 0, 150)
    rg, _ = gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, 150, FP64)
    compare_chains(ro, rg)
    gm = workloads.load(gpu, sysm)
    rm, _ = gpu.bisection_run(gm, sysm.num_levels, sysm.seed, 0, 150, MIXED)
    worst, n = 0.0, 0
    for r in range(2):
        for m in range(150):
            ns = ro["num_stages"][r, m]
            if ns != rm["num_stages"][r, m] or ro["accepted"][r, m] != rm["accepted"][r, m]:
                break
            d = (ro["new_action"][r, m, :ns] - ro["old_action"][r, m, :ns]) - \
                (rm["new_action"][r, m, :ns] - rm["old_action"][r, m, :ns])
            worst = max(worst, float(np.max(np.abs(d))))
            n += ns
    assert n > 100 and worst <= 1e-6


def test_bc_fit_without_ulong_is_rejected(gpu):
    qg, tg = (abi.GRID_LOG, 0.01, 6.0), (abi.GRID_LINEAR, 0.0, 1.0)
    umat = smooth_umat(3, tables._grid(*qg, 30), tables._grid(*tg, 8))
    fid = gpu.fit_create_coulomb_bc(qg, tg, umat, 0.125, -1.0)
    rid = gpu.replica_create(4, 8, (5.0, 5.0, 5.0), 0.125, [4], [0.5], [fid], use_long_range=True)
    gpu.path_set(rid, np.random.default_rng(0).uniform(-2, 2, (9, 4, 3)))
    with pytest.raises(RuntimeError):
        gpu.action_short_range([(rid, 0, 0, 8, 0, [1])])
    start, end, coefs = ewald.ulong_quintic(-1.0, 0.5, 0.125, 3, 0.75 * math.sqrt(75.0))
    gpu.set_ulong(rid, 0, 0, start, end, coefs)
    assert np.isfinite(gpu.action_short_range([(rid, 0, 0, 8, 0, [1])])[0])
