"""Permutation bookkeeping on the GPU against the oracle: HTable
(PermuteTableClass::ConstructHTable) to float rounding, CycleClass::Apply and the
accept/reject copies of the permuted particles bit-exact."""

import numpy as np
import pytest

from pimcplusplus_amd import abi, workloads
from tests.test_gpu_parity import compare_chains
from tests.test_oracle_permute import two_species

pytestmark = pytest.mark.gpu


def test_htable_matches_oracle_on_a_sampled_path(oracle, gpu):
    sysm = workloads.config0(num_replicas=1)
    io, ig = workloads.load(oracle, sysm), workloads.load(gpu, sysm)
    print("config0 path after 60 moves", flush=True)
    oracle.bisection_run(io, sysm.num_levels, sysm.seed, 0, 60)
    gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, 60, abi.FP64)
    n = sysm.replicas[0].num_particles
    for s1, s2, eps in ((0, 8, 1e-6), (5, 21, 1e-3), (30, 32, 1e-10)):
        ho = oracle.permute_htable(io[0], 0, s1, s2, eps, n)
        hg = gpu.permute_htable(ig[0], 0, s1, s2, eps, n)
        assert np.array_equal(ho == 0, hg == 0)
        nz = ho != 0
        assert np.max(np.abs(hg[nz] - ho[nz]) / ho[nz]) <= 1.2e-7      # one float rounding
        assert np.mean(hg == ho) > 0.99


def test_htable_two_species_and_modes(oracle, gpu):
    io, beads = two_species(oracle)
    ig, _ = two_species(gpu)
    moved = beads.copy()
    moved[5, 9] += 0.25
    for lib, r in ((oracle, io), (gpu, ig)):
        lib.path_set(r, moved, 1)
    for species, n in ((0, 7), (1, 5)):
        for mode in (0, 1):
            ho = oracle.permute_htable(io, species, 1, 9, 1e-5, n, mode)
            hg = gpu.permute_htable(ig, species, 1, 9, 1e-5, n, mode)
            assert np.array_equal(ho == 0, hg == 0)
            assert np.allclose(hg, ho, rtol=1.2e-7, atol=0)


def test_cycle_apply_and_copies_bit_exact(oracle, gpu):
    io, _ = two_species(oracle)
    ig, _ = two_species(gpu)
    rng = np.random.default_rng(5)
    for step in range(12):
        first, n = (0, 7) if step % 2 == 0 else (7, 5)
        length = 1 + step % 4
        cycle = [int(c) for c in rng.choice(n, size=length, replace=False)]
        sl = int(rng.integers(0, 17))
        for lib, r in ((oracle, io), (gpu, ig)):
            lib.cycle_apply(r, first, sl, cycle)
        for mode in (0, 1):
            assert np.array_equal(oracle.path_get(io, mode), gpu.path_get(ig, mode)), (step, mode)
            assert np.array_equal(oracle.permutation(io, mode), gpu.permutation(ig, mode)), (step, mode)
        ptcls = [first + c for c in cycle]
        for lib, r in ((oracle, io), (gpu, ig)):
            if step % 3 == 2:
                lib.reject_copy(r, sl, sl, ptcls)
            else:
                lib.accept_copy(r, sl, sl, ptcls)
        for mode in (0, 1):
            assert np.array_equal(oracle.path_get(io, mode), gpu.path_get(ig, mode)), (step, mode)
            assert np.array_equal(oracle.permutation(io, mode), gpu.permutation(ig, mode)), (step, mode)


@pytest.mark.parametrize("joins", [(0, 7, 2), (3, 11, 12, 5, 0), (12, 1, 9)])
def test_move_join_permuted_world_lines_bit_exact(oracle, gpu, joins):
    from tests.test_oracle_join import permuted_replica, world_line_storage
    ro, perm = permuted_replica(oracle)
    rg, pg = permuted_replica(gpu)
    assert np.array_equal(perm, pg)
    W = np.random.default_rng(sum(joins)).uniform(-3.0, 3.0, (9, 13, 3))
    for lib, r in ((oracle, ro), (gpu, rg)):
        lib.path_set(r, world_line_storage(W, perm, joins[0]), abi.BOTHMODES)
    for old_j, new_j in zip(joins[:-1], joins[1:]):
        print("move_join", old_j, new_j, flush=True)
        oracle.move_join(ro, old_j, new_j)
        gpu.move_join(rg, old_j, new_j)
        want = world_line_storage(W, perm, new_j)
        for mode in (abi.OLDMODE, abi.NEWMODE):
            assert np.array_equal(gpu.path_get(rg, mode), want), (old_j, new_j, mode)
            assert np.array_equal(gpu.path_get(rg, mode), oracle.path_get(ro, mode))
        assert gpu.join(rg) == oracle.join(ro) == new_j


def test_move_join_unequal_copies_bit_exact(oracle, gpu):
    from tests.test_oracle_join import permuted_replica
    ro, perm = permuted_replica(oracle)
    rg, _ = permuted_replica(gpu)
    rng = np.random.default_rng(5)
    for old_j, new_j in ((2, 9), (9, 1), (1, 1), (0, 12)):
        old, new = rng.normal(size=(13, 9, 3)), rng.normal(size=(13, 9, 3))
        for lib, r in ((oracle, ro), (gpu, rg)):
            lib.path_set(r, old, abi.OLDMODE)
            lib.path_set(r, new, abi.NEWMODE)
            lib.move_join(r, old_j, new_j)
        for mode in (abi.OLDMODE, abi.NEWMODE):
            assert np.array_equal(gpu.path_get(rg, mode), oracle.path_get(ro, mode)), (old_j, new_j, mode)


@pytest.mark.parametrize("L", [2.0, 0.5, 3.0, 7.1529])
def test_put_in_box_bit_exact_with_ties(oracle, gpu, L):
    y = np.zeros((2, 10, 7))
    ids = []
    for lib in (oracle, gpu):
        fid = lib.fit_create_david(0, 0.1, L / 2, y, 2, 0.025, 0.025)
        ids.append(lib.replica_create(1, 2, (L, L, L), 0.025, [1], [6.0597], [fid]))
    ties = (np.arange(-4, 4) + 0.5) * L
    rng = np.random.default_rng(int(L * 100))
    v = np.concatenate([np.stack([ties, -ties, ties[::-1]], axis=1), rng.uniform(-6 * L, 6 * L, (400, 3)),
                        np.array([[L / 2, -L / 2, 0.0]])])
    for fast in (False, True):
        a = oracle.put_in_box(ids[0], v, fast=fast)
        b = gpu.put_in_box(ids[1], v, fast=fast)
        assert np.array_equal(a, b), fast
    if L in (2.0, 0.5):          # exact ties: PutInBox sends +-L/2 to -L/2
        assert np.all(gpu.put_in_box(ids[1], v[:8]) == -L / 2)


def permute_config0(lib, sysm, ids, cycles=((0, 5, 9), (2, 11, 7, 3))):
    """Cycles applied at slice 8 of every replica (CycleClass::Apply) and accepted."""
    for r in ids:
        for c in cycles:
            lib.cycle_apply(r, 0, 8, list(c))
            lib.accept_copy(r, 8, 8, list(c))


def test_chain_with_permuted_replicas_matches_oracle(oracle, gpu):
    from tests.test_gpu_parity import compare_chains
    sysm = workloads.config0(num_replicas=3)
    io, ig = workloads.load(oracle, sysm), workloads.load(gpu, sysm)
    permute_config0(oracle, sysm, io[:2])       # replica 2 keeps the identity permutation
    permute_config0(gpu, sysm, ig[:2])
    for first in (0, 150):
        print("permuted chain, moves", first, flush=True)
        ro, so = oracle.bisection_run(io, sysm.num_levels, sysm.seed, first, 150)
        rg, sg = gpu.bisection_run(ig, sysm.num_levels, sysm.seed, first, 150, abi.FP64)
        compare_chains(ro, rg)
        assert (so.moves, so.accepted, so.stages) == (sg.moves, sg.accepted, sg.stages)
    for a, b in zip(io, ig):
        assert oracle.join(a) == gpu.join(b)
        assert np.array_equal(oracle.permutation(a), gpu.permutation(b))
        for mode in (abi.OLDMODE, abi.NEWMODE):
            assert np.max(np.abs(oracle.path_get(a, mode) - gpu.path_get(b, mode))) < 1e-12
    assert not np.array_equal(gpu.permutation(ig[0]), np.arange(sysm.replicas[0].num_particles))


def test_cluster_plan_refuses_a_permuted_replica(gpu):
    sysm = workloads.config0(num_replicas=2)
    ig = workloads.load(gpu, sysm)
    permute_config0(gpu, sysm, ig[:1])
    gpu.set_plan(256, 1, False, 2)
    with pytest.raises(abi.PimcError, match="permutation"):
        gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, 4, abi.FP64)
    gpu.set_plan(256, 1, False, 2)
    gpu.bisection_run(ig[1:], sysm.num_levels, sysm.seed, 0, 4, abi.FP64)   # identity: runs
    plan = gpu.bisection_tune(ig, sysm.num_levels, sysm.seed, 0, 2, abi.FP64)
    assert plan.cluster == 1


def table_chains(lib, sysm, nmoves, precision=abi.FP64, gamma=(1.0, 1.0, 1.0, 1.0), eps=1e-4):
    ids = workloads.load(lib, sysm)
    for r in ids:
        lib.set_permute(r, 0, gamma, eps)
    rec, st = lib.bisection_run(ids, sysm.num_levels, sysm.seed, 0, nmoves, precision)
    return ids, rec, st


def test_table_permutation_chain_matches_the_oracle(oracle, gpu):
    """PermuteType TABLE moves in the resident chain: the drawn cycles, the bisection stages of
    every cycle particle, the final permutation test and the accepted beads and permutation
    entries equal the oracle's (the cycle weights are float, so forward_t and reverse_t agree
    to float rounding; the rest as the single-particle chain)."""
    sysm = workloads.dilute_boson_system(num_replicas=3)
    print("dilute bosons, TABLE, FP64, 3 x 300 moves", flush=True)
    io, ro, so = table_chains(oracle, sysm, 300)
    ig, rg, sg = table_chains(gpu, sysm, 300)
    assert np.array_equal(ro["cycle_length"], rg["cycle_length"])
    assert np.array_equal(ro["cycle"], rg["cycle"])
    assert np.array_equal(ro["num_cycles"], rg["num_cycles"])
    compare_chains(ro, rg)
    for key in ("forward_t", "reverse_t", "accept_prob"):
        a, b = ro[key], rg[key]
        assert np.max(np.abs(a - b) / (1e-30 + np.abs(a))) < 1e-6, key
    assert np.array_equal(ro["psi"], rg["psi"])
    acc = ro["accepted"].astype(bool)
    assert np.sum(acc & (ro["cycle_length"] > 1)) >= 3      # exchanges were accepted
    assert (so.moves, so.accepted, so.stages, so.pair_evals) == (sg.moves, sg.accepted, sg.stages, sg.pair_evals)
    for a, b in zip(io, ig):
        assert np.max(np.abs(oracle.path_get(a) - gpu.path_get(b, 0))) < 1e-12
        assert np.array_equal(gpu.path_get(b, 0), gpu.path_get(b, 1))
        assert np.array_equal(oracle.permutation(a, 0), gpu.permutation(b, 0))
        assert np.array_equal(gpu.permutation(b, 0), gpu.permutation(b, 1))
        assert oracle.join(a) == gpu.join(b)


@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_table_permutation_chain_generic_kernel_matches_the_oracle(oracle, gpu, precision):
    """The generic kernel's TABLE moves (a bicubic Coulomb pair action on a log grid), where
    2-, 3- and 4-cycles are drawn and accepted: FP64 chains equal the oracle's, the mixed
    chain's stage changes stay within 1e-6 of them."""
    sysm = workloads.coulomb_boson_system(num_replicas=2)
    print("Coulomb bosons, TABLE, generic kernel, 2 x 250 moves, precision", precision, flush=True)
    io, ro, so = table_chains(oracle, sysm, 250)
    ig, rg, sg = table_chains(gpu, sysm, 250, precision)
    cl, acc = ro["cycle_length"], ro["accepted"].astype(bool)
    assert all(np.sum(acc & (cl == k)) >= 2 for k in (2, 3, 4))
    if precision == abi.FP64:
        assert np.array_equal(ro["cycle"], rg["cycle"])
        compare_chains(ro, rg)
        for key in ("forward_t", "reverse_t", "accept_prob"):
            assert np.max(np.abs(ro[key] - rg[key]) / (1e-30 + np.abs(ro[key]))) < 1e-6, key
        for a, b in zip(io, ig):
            assert np.max(np.abs(oracle.path_get(a) - gpu.path_get(b, 0))) < 1e-12
            assert np.array_equal(oracle.permutation(a, 0), gpu.permutation(b, 0))
        return
    worst, n = 0.0, 0
    for r in range(2):
        for m in range(250):
            ns = ro["num_stages"][r, m]
            if ns != rg["num_stages"][r, m] or ro["accepted"][r, m] != rg["accepted"][r, m] or \
                    ro["cycle_length"][r, m] != rg["cycle_length"][r, m]:
                break
            d = (ro["new_action"][r, m, :ns] - ro["old_action"][r, m, :ns]) - \
                (rg["new_action"][r, m, :ns] - rg["old_action"][r, m, :ns])
            worst = max(worst, float(np.max(np.abs(d))))
            n += ns
    assert n > 100
    assert worst <= 1e-6


def test_table_permutation_chain_mixed_within_the_bound(oracle, gpu):
    sysm = workloads.dilute_boson_system(num_replicas=2)
    print("dilute bosons, TABLE, mixed, 2 x 200 moves", flush=True)
    _, ro, _ = table_chains(oracle, sysm, 200)
    _, rg, _ = table_chains(gpu, sysm, 200, abi.MIXED)
    worst, n = 0.0, 0
    for r in range(2):
        for m in range(200):
            ns = ro["num_stages"][r, m]
            if ns != rg["num_stages"][r, m] or ro["accepted"][r, m] != rg["accepted"][r, m] or \
                    ro["cycle_length"][r, m] != rg["cycle_length"][r, m]:
                break
            d = (ro["new_action"][r, m, :ns] - ro["old_action"][r, m, :ns]) - \
                (rg["new_action"][r, m, :ns] - rg["old_action"][r, m, :ns])
            worst = max(worst, float(np.max(np.abs(d))))
            n += ns
    assert n > 100
    assert worst <= 1e-6


def test_table_permutation_moves_refuse_unsupported_launches(gpu):
    sysm = workloads.dilute_boson_system(num_replicas=2)
    ids = workloads.load(gpu, sysm)
    gpu.set_permute(ids[0], 0, (1.0, 1.0, 1.0, 1.0), 1e-4)
    with pytest.raises(abi.PimcError) as e:       # one replica with TABLE, one without
        gpu.bisection_run(ids, sysm.num_levels, sysm.seed, 0, 1, abi.FP64)
    assert e.value.code == abi.ERR_UNSUPPORTED
    gpu.set_permute(ids[1], 0, (1.0, 1.0, 1.0, 1.0), 1e-4)
    with pytest.raises(abi.PimcError) as e:
        gpu.sweep_graph_create(ids, sysm.num_levels, sysm.seed, 0, 2, abi.FP64)
    assert e.value.code == abi.ERR_UNSUPPORTED
    gpu.set_plan(256, 1, False, 2)
    with pytest.raises(abi.PimcError) as e:
        gpu.bisection_run(ids, sysm.num_levels, sysm.seed, 0, 1, abi.FP64)
    assert e.value.code == abi.ERR_UNSUPPORTED
    gpu.set_plan(256, 1, False, 1)
    gpu.set_permute(ids[0], 0)                    # back to no permutations
    gpu.set_permute(ids[1], 0)
    gpu.bisection_run(ids, sysm.num_levels, sysm.seed, 0, 2, abi.FP64)


def test_table_permutation_tuner_tries_link_splits_and_restores_the_chain(oracle, gpu):
    """The tuner times TABLE moves with link splits 1, 2 and 4 (one 256-thread block per
    replica), restores the replicas after every trial, and the chain that follows with the
    chosen plan equals the oracle's (FP64: only the order of the partner sums changes)."""
    sysm = workloads.dilute_boson_system(num_replicas=3)
    print("dilute bosons, TABLE, tuned plan, 3 x 200 moves", flush=True)
    ig = workloads.load(gpu, sysm)
    for r in ig:
        gpu.set_permute(r, 0, (1.0, 1.0, 1.0, 1.0), 1e-4)
    before = [gpu.path_get(r, 0) for r in ig]
    plan = gpu.bisection_tune(ig, sysm.num_levels, sysm.seed, 0, 8, abi.FP64)
    assert (plan.block_threads, plan.persistent, plan.cluster) == (256, 0, 1) and plan.link_split in (1, 2, 4)
    assert all(np.array_equal(b, gpu.path_get(r, 0)) for b, r in zip(before, ig))
    rg, sg = gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, 200, abi.FP64)
    io, ro, so = table_chains(oracle, sysm, 200)
    assert np.array_equal(ro["cycle"], rg["cycle"])
    compare_chains(ro, rg)
    for link_split in (2, 4):   # every split the tuner may pick gives the oracle's chain
        gs = workloads.load(gpu, sysm)
        for r in gs:
            gpu.set_permute(r, 0, (1.0, 1.0, 1.0, 1.0), 1e-4)
        gpu.set_plan(256, link_split, False, 1)
        rs, _ = gpu.bisection_run(gs, sysm.num_levels, sysm.seed, 0, 200, abi.FP64)
        assert np.array_equal(ro["cycle"], rs["cycle"])
        compare_chains(ro, rs)
    gpu.set_plan(256, 1, False, 1)


@pytest.mark.parametrize("n,levels,moves,lo,hi", [(20, 3, 40, 5461, 65536), (32, 4, 12, 65536, 1 << 30)])
def test_table_permutation_large_cycle_tables_match_the_oracle(oracle, gpu, n, levels, moves, lo, hi):
    """Helium atoms at the liquid density: 20 atoms with 8-slice windows give cycle tables of
    ~8e3 entries, beyond a block's 64 KB of LDS (5461 entries at most), which the block stores
    in its global scratch; 32 atoms with 16-slice windows give ~2e5 entries, beyond that scratch
    too, so the kernel walks the reference's loops with one thread.  The chains still equal the
    oracle's."""
    sysm = workloads.lj_system("dense-bosons", n, 32, 0.05, levels, 1234, num_replicas=2)
    print(f"dense bosons, TABLE, 2 x {moves} moves, {n} atoms, level {levels}", flush=True)
    io, ro, so = table_chains(oracle, sysm, moves)
    ig, rg, sg = table_chains(gpu, sysm, moves)
    assert lo < ro["num_cycles"].min() and ro["num_cycles"].max() <= hi
    assert np.array_equal(ro["num_cycles"], rg["num_cycles"])
    assert np.array_equal(ro["cycle"], rg["cycle"])
    compare_chains(ro, rg)
    for key in ("forward_t", "reverse_t", "accept_prob"):
        assert np.max(np.abs(ro[key] - rg[key]) / (1e-30 + np.abs(ro[key]))) < 1e-6, key
    for a, b in zip(io, ig):
        assert np.max(np.abs(oracle.path_get(a) - gpu.path_get(b, 0))) < 1e-12
        assert np.array_equal(oracle.permutation(a, 0), gpu.permutation(b, 0))
