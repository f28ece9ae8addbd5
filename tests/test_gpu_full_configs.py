"""Parity at the benchmark configurations' full sizes.

The oracle runs a few replicas of each configuration at its real N, M and bisection
level; the GPU runs the same replicas and the chains must agree move by move
(accept/reject sequence identical, actions to 1e-12).  At the benchmark's full
replica count (4096 replicas of N = 256, M = 64) the oracle checks a sample of
replicas, and the size-independent property is that every launch shape gives the same
chains bit for bit."""

import os

import numpy as np
import pytest

from pimcplusplus_amd import abi, workloads
from pimcplusplus_amd.tables import load_npz
from tests.test_gpu_parity import compare_chains

pytestmark = pytest.mark.gpu


def _mixed_delta_s_error(ro, rm, nrep, nmoves):
    """max |dS_mixed - dS_oracle| over the stages both chains attempted before they first differ."""
    worst, n = 0.0, 0
    for r in range(nrep):
        for m in range(nmoves):
            ns = ro["num_stages"][r, m]
            if ns != rm["num_stages"][r, m] or ro["accepted"][r, m] != rm["accepted"][r, m]:
                break
            d = (ro["new_action"][r, m, :ns] - ro["old_action"][r, m, :ns]) - \
                (rm["new_action"][r, m, :ns] - rm["old_action"][r, m, :ns])
            worst = max(worst, float(np.max(np.abs(d))))
            n += ns
    return worst, n


@pytest.mark.parametrize("m", [128, 168])
def test_config2_hydrogen_full_size(oracle, gpu, m):
    # M = 128 as stated (T = 19,737 K at tau = 0.125) and M = 168 (T = 15,037 K, the stated temperature)
    sysm = workloads.config2(load_npz(os.path.join(os.path.dirname(__file__), "golden", "testcoul_fit2.npz")),
                             num_replicas=2, m=m)
    print(f"config2: 32 e + 32 p, M={m}, T={workloads.temperature_k(sysm):.0f} K, level 4, breakup tables", flush=True)
    io, ig = workloads.load(oracle, sysm), workloads.load(gpu, sysm)
    ro, so = oracle.bisection_run(io, sysm.num_levels, sysm.seed, 0, 150)
    rg, sg = gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, 150, abi.FP64)
    compare_chains(ro, rg)
    assert so.k_evals == sg.k_evals > 0
    for a, b in zip(io, ig):
        assert np.max(np.abs(oracle.rho_k(a) - gpu.rho_k(b))) < 1e-10
    gm = workloads.load(gpu, sysm)
    rm, _ = gpu.bisection_run(gm, sysm.num_levels, sysm.seed, 0, 150, abi.MIXED)
    worst, n = _mixed_delta_s_error(ro, rm, 2, 150)
    print(f"config2 mixed: max |dS error| {worst:.3g} over {n} stages", flush=True)
    assert n > 200 and worst <= 1e-6


@pytest.mark.parametrize("cluster", [1, 16])
def test_config4_ocp_full_size(oracle, gpu, cluster):
    sysm = workloads.config4(num_replicas=2)
    print(f"config4: OCP N=1024, M=128, level 4, 2069 k, cluster={cluster}", flush=True)
    io = workloads.load(oracle, sysm)
    gpu.set_plan(256, 4 if cluster > 1 else 1, False, cluster)
    ig = workloads.load(gpu, sysm)
    nmoves = 150
    ro, so = oracle.bisection_run(io, sysm.num_levels, sysm.seed, 0, nmoves)
    rg, sg = gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, nmoves, abi.FP64)
    compare_chains(ro, rg)
    assert (so.accepted, so.k_evals) == (sg.accepted, sg.k_evals)
    for a, b in zip(io, ig):
        assert np.max(np.abs(oracle.path_get(a) - gpu.path_get(b))) < 1e-12
        assert np.max(np.abs(oracle.rho_k(a) - gpu.rho_k(b))) < 1e-9
    # the mixed path at this size (FP32 bicubic, Ulong tail and k-space phase sums)
    gm = workloads.load(gpu, sysm)
    rm, _ = gpu.bisection_run(gm, sysm.num_levels, sysm.seed, 0, nmoves, abi.MIXED)
    worst, n = _mixed_delta_s_error(ro, rm, 2, nmoves)
    print(f"config4 mixed: max |dS error| {worst:.3g} over {n} stages", flush=True)
    assert n > 200 and worst <= 1e-6


def test_config3_uneven_full_size(oracle, gpu):
    sysm = workloads.config3(num_replicas=128)
    print("config3: 128 uneven LJ replicas", flush=True)
    io = workloads.load(oracle, sysm)
    gpu.set_plan(512, 1, True)
    ig = workloads.load(gpu, sysm)
    ro, so = oracle.bisection_run(io, sysm.num_levels, sysm.seed, 0, 10)
    rg, sg = gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, 10, abi.FP64)
    compare_chains(ro, rg)
    assert (so.moves, so.accepted, so.pair_evals) == (sg.moves, sg.accepted, sg.pair_evals)


def test_config1_benchmark_size_launch_shapes_agree_and_sample_matches_oracle(oracle):
    from pimcplusplus_amd import native
    sysm = workloads.config1(num_replicas=4096)
    print("config1: 4096 replicas, 8 moves, two launch shapes", flush=True)
    paths = []
    for plan in ((256, 1, False, 1), (256, 1, True, 1)):
        lib = native.open_context(0)
        try:
            lib.set_plan(*plan)
            ids = workloads.load(lib, sysm)
            lib.bisection_run(ids, sysm.num_levels, sysm.seed, 0, 8, abi.FP64, records=False)
            sample = [0, 1, 1000, 2047, 3001, 4095]
            paths.append([lib.path_get(ids[i]) for i in sample])
            st = lib.bisection_stats_take(ids)
            assert st.moves == 8 * 4096
        finally:
            lib.close()
    for a, b in zip(*paths):
        assert np.array_equal(a, b)
    # the oracle reruns the sampled replicas (each replica's chain depends on its own stream only)
    for i, p in zip(sample, paths[0]):
        spec = sysm.replicas[i]
        one = workloads.System(name="one", tables=sysm.tables, desired_tau=sysm.desired_tau, replicas=[spec],
                               num_levels=sysm.num_levels, seed=sysm.seed)
        rid = workloads.load(oracle, one)
        oracle.bisection_run(rid, sysm.num_levels, sysm.seed, 0, 8)
        assert np.max(np.abs(oracle.path_get(rid[0]) - p)) < 1e-12


def test_config1_ten_thousand_moves_against_the_oracle(oracle, gpu):
    """The benchmark system over 10^4 seeded moves (4 replicas x 2500): the FP64 chain equals
    the oracle's move by move, and the mixed chain's dS stays within 1e-6 of it."""
    nrep, nmoves = 4, 2500
    sysm = workloads.config1(num_replicas=nrep)
    print(f"config1: {nrep} replicas x {nmoves} moves, FP64 and mixed against the oracle", flush=True)
    io, ig = workloads.load(oracle, sysm), workloads.load(gpu, sysm)
    ro, so = oracle.bisection_run(io, sysm.num_levels, sysm.seed, 0, nmoves)
    rg, sg = gpu.bisection_run(ig, sysm.num_levels, sysm.seed, 0, nmoves, abi.FP64)
    compare_chains(ro, rg)
    assert (so.moves, so.accepted, so.pair_evals) == (sg.moves, sg.accepted, sg.pair_evals) and so.moves == nrep * nmoves
    gm = workloads.load(gpu, sysm)
    rm, _ = gpu.bisection_run(gm, sysm.num_levels, sysm.seed, 0, nmoves, abi.MIXED)
    worst, n = _mixed_delta_s_error(ro, rm, nrep, nmoves)
    print(f"mixed: max |dS error| {worst:.3g} over {n} stages", flush=True)
    assert n > 1000 and worst <= 1e-6
