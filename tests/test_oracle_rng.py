"""Pins the oracle's random stream: Philox4x32-10 known answers, the uniform
mapping, and the polar Box-Muller pairs."""

import ctypes

import numpy as np

from tests.helpers import oracle_cdll


def philox(ctr, key):
    lib = oracle_cdll()
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib.oracle_philox4x32_10(c, k, o)
    return [o[i] for i in range(4)]


def test_philox_known_answers():
    # Known-answer vectors of Philox4x32-10 published with Random123 (Salmon et al., SC'11).
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_mapping():
    lib = oracle_cdll()
    assert lib.oracle_uniform(0, 0) == 0.0
    assert lib.oracle_uniform(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    # the high word carries the top bits: w1 = 2^31 -> 0.5
    assert lib.oracle_uniform(0, 0x80000000) == 0.5
    assert lib.oracle_uniform(0x800, 0) == 2.0 ** -53


def test_gaussian_pairs_are_standard_normal():
    lib = oracle_cdll()
    g = (ctypes.c_double * 2)()
    vals = []
    attempts = []
    for p in range(20000):
        attempts.append(lib.oracle_gaussian_pair(1234, 7, 3, 1, p, g))
        vals += [g[0], g[1]]
    v = np.array(vals)
    n = v.size
    assert abs(v.mean()) < 4.0 / np.sqrt(n)
    assert abs(v.var() - 1.0) < 6.0 * np.sqrt(2.0 / n)
    # fourth moment of a standard normal is 3
    assert abs(np.mean(v ** 4) - 3.0) < 0.15
    # polar method: acceptance pi/4, so about 4/pi attempts per pair
    assert abs(np.mean(attempts) - 4.0 / np.pi) < 0.02
    # the two variates of a pair are uncorrelated
    assert abs(np.corrcoef(v[0::2], v[1::2])[0, 1]) < 0.03


def test_gaussian_pair_reproduces_polar_method_from_philox_words():
    lib = oracle_cdll()
    seed, stream, move, level, pair = 0x0123456789ABCDEF, 5, 77, 2, 9
    g = (ctypes.c_double * 2)()
    used = lib.oracle_gaussian_pair(seed, stream, move, level, pair, g)
    # replay with the documented counter layout
    for a in range(used):
        w = philox([move, (3 << 24) | (level << 16) | a, pair, stream], [seed & 0xFFFFFFFF, seed >> 32])
        v1 = 2.0 * lib.oracle_uniform(w[0], w[1]) - 1.0
        v2 = 2.0 * lib.oracle_uniform(w[2], w[3]) - 1.0
        s = v1 * v1 + v2 * v2
        if 0.0 < s < 1.0:
            assert a == used - 1
            t = np.sqrt(-2.0 * np.log(s) / s)
            assert g[0] == v1 * t and g[1] == v2 * t
