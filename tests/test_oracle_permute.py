"""Pins the oracle's permutation bookkeeping: PermuteTableClass::ConstructHTable
(PermuteTableClass.cc:122-150) and CycleClass::Apply (PermuteTableClass.cc:172-187).

HTable is checked against a brute-force nearest-image computation in numpy (images
-2..2 in every dimension) and its closed forms: the diagonal is exp(0) = 1, a path
that does not move between the two slices gives the symmetric free-particle
propagator ratio exp(-|r_j - r_i|^2 / 4 lambda beta), and entries below epsilon are
exactly zero.  Apply is checked against numpy index rolls."""

import itertools

import numpy as np
import pytest

LAMBDA, TAU = 6.0597, 0.025


def two_species(lib, n0=7, n1=5, m=16, box=8.0, seed=3, static=False):
    y = np.zeros((4, 10, 7))
    fid = lib.fit_create_david(0, 0.1, box / 2, y, 2, TAU, TAU)
    n = n0 + n1
    rid = lib.replica_create(n, m, (box,) * 3, TAU, [n0, n1], [LAMBDA, 2 * LAMBDA], [fid] * 4, rng_stream=1)
    rng = np.random.default_rng(seed)
    beads = rng.uniform(-box / 2, box / 2, (m + 1, n, 3))
    if static:
        beads[:] = beads[0]
    lib.path_set(rid, beads)
    return rid, beads


def nearest(d, box):
    """Brute-force nearest image of displacement rows d over images -2..2."""
    best = np.full(d.shape[:-1], np.inf)
    out = np.zeros_like(d)
    for img in itertools.product(range(-2, 3), repeat=3):
        e = d + box * np.array(img)
        r2 = np.sum(e * e, axis=-1)
        take = r2 < best
        best = np.where(take, r2, best)
        out = np.where(take[..., None], e, out)
    return out


def brute_htable(beads, first, n, s1, s2, lam, box, eps):
    a = beads[s1, first:first + n]
    b = beads[s2, first:first + n]
    dii = np.sum(nearest(b - a, box) ** 2, axis=-1)
    dij = np.sum(nearest(b[None, :, :] - a[:, None, :], box) ** 2, axis=-1)
    x = (-dij + dii[:, None]) / (4.0 * lam * TAU * (s2 - s1))
    return np.where(x < np.log(eps), 0.0, np.exp(x))


@pytest.mark.parametrize("species,first,n,lam", [(0, 0, 7, LAMBDA), (1, 7, 5, 2 * LAMBDA)])
def test_htable_matches_brute_force(oracle, species, first, n, lam):
    rid, beads = two_species(oracle)
    for s1, s2, eps in ((0, 16, 1e-4), (3, 11, 1e-2), (8, 9, 1e-30)):
        h = oracle.permute_htable(rid, species, s1, s2, eps, n)
        want = brute_htable(beads, first, n, s1, s2, lam, 8.0, eps)
        assert np.all((h == 0) == (want == 0))
        assert np.allclose(h, want, rtol=2e-7, atol=0)
        assert np.all(np.diag(h) == 1.0)


def test_htable_of_a_static_path_is_the_symmetric_propagator_ratio(oracle):
    rid, beads = two_species(oracle, static=True)
    h = oracle.permute_htable(rid, 0, 2, 10, 1e-6, 7)
    r = beads[0, :7]
    d = nearest(r[None, :, :] - r[:, None, :], 8.0)
    want = np.exp(-np.sum(d * d, axis=-1) / (4 * LAMBDA * TAU * 8))
    want[want < 1e-6] = 0.0
    assert np.array_equal(h, h.T)
    assert np.allclose(h, want, rtol=2e-7, atol=1e-12)


def test_htable_reads_the_requested_copy(oracle):
    rid, beads = two_species(oracle)
    moved = beads.copy()
    moved[5, 2] += 0.3
    oracle.path_set(rid, moved, 1)
    h0 = oracle.permute_htable(rid, 0, 0, 5, 1e-8, 7, mode=0)
    h1 = oracle.permute_htable(rid, 0, 0, 5, 1e-8, 7, mode=1)
    assert np.allclose(h0, brute_htable(beads, 0, 7, 0, 5, LAMBDA, 8.0, 1e-8), rtol=2e-7, atol=0)
    assert np.allclose(h1, brute_htable(moved, 0, 7, 0, 5, LAMBDA, 8.0, 1e-8), rtol=2e-7, atol=0)
    assert not np.array_equal(h0[:, 2], h1[:, 2])


@pytest.mark.parametrize("cycle", [[3], [0, 4], [4, 1, 2], [2, 4, 0, 3]])
def test_cycle_apply_is_an_index_roll(oracle, cycle):
    rid, beads = two_species(oracle)
    perm0 = oracle.permutation(rid, 1)
    oracle.cycle_apply(rid, 7, 9, cycle)          # species 1 starts at particle 7; slice 9
    new, old = oracle.path_get(rid, 1), oracle.path_get(rid, 0)
    idx = [7 + c for c in cycle]
    want = beads.copy()
    want[9, idx] = beads[9, np.roll(idx, -1)]
    assert np.array_equal(new, want)
    assert np.array_equal(old, beads)
    p = oracle.permutation(rid, 1)
    wp = perm0.copy()
    wp[idx] = perm0[np.roll(idx, -1)]
    assert np.array_equal(p, wp)
    assert np.array_equal(oracle.permutation(rid, 0), perm0)
    for _ in range(len(cycle) - 1):              # a cycle of length L has order L
        oracle.cycle_apply(rid, 7, 9, cycle)
    assert np.array_equal(oracle.path_get(rid, 1), beads)
    assert np.array_equal(oracle.permutation(rid, 1), perm0)


@pytest.mark.parametrize("first,cycle", [(0, [1, 1]), (7, [5]), (0, [0, 1, 2, 3, 4]), (0, []), (-1, [0])])
def test_cycle_apply_rejects_bad_cycles(oracle, first, cycle):
    rid, _ = two_species(oracle)
    with pytest.raises(RuntimeError):
        oracle.cycle_apply(rid, first, 3, cycle)
