"""Pins of the oracle's MRG32k3a (random/MRG.hpp, random/src/MRG32k3a.cpp).

- the published known answer of L'Ecuyer's generator for the default seed
  (12345 x 6): first output 0.127011122046577 (RngStreams test output);
- the reference's own jump matrices A^(2^76), A^(2^127) (golden fixture from
  random/MRG.hpp:50-69) against the oracle's repeated squaring;
- substream i = (A^(2^76))^i seed, by exact big-integer matrix arithmetic;
- the recurrence against the one-step transition matrices A1p0/A2p0
  (MRG.hpp:71-75) applied n times, in Python integers.
"""
import json
import os

import numpy as np

import oracle_harness as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_constants.json")))

M1, M2 = 4294967087, 4294944443
A1 = [[0, 1, 0], [0, 0, 1], [-810728, 1403580, 0]]
A2 = [[0, 1, 0], [0, 0, 1], [-1370589, 0, 527612]]


def matmul(a, b, m):
    return [[sum(a[i][k] * b[k][j] for k in range(3)) % m for j in range(3)] for i in range(3)]


def matpow(a, e, m):
    r = [[int(i == j) for j in range(3)] for i in range(3)]
    b = [[x % m for x in row] for row in a]
    while e:
        if e & 1:
            r = matmul(r, b, m)
        b = matmul(b, b, m)
        e >>= 1
    return r


def matvec(a, v, m):
    return [sum(a[i][k] * v[k] for k in range(3)) % m for i in range(3)]


def test_known_answer_default_seed():
    st = O.rng_init(np.full(6, 12345, dtype=np.uint32), [0])
    d, _ = O.rng_draw(st, 3)
    assert abs(d[0, 0] - 0.127011122046577) < 1e-15


def test_jump_matrices_match_reference_constants():
    J = O.jump_matrices()
    assert np.array_equal(J[0, 0], np.array(GOLD["A1p76"], dtype=np.uint64))
    assert np.array_equal(J[1, 0], np.array(GOLD["A2p76"], dtype=np.uint64))
    assert np.array_equal(J[0, 51], np.array(GOLD["A1p127"], dtype=np.uint64))
    assert np.array_equal(J[1, 51], np.array(GOLD["A2p127"], dtype=np.uint64))
    # every level is the square of the previous one
    for c, m in ((0, M1), (1, M2)):
        for j in (1, 7, 30, 63):
            prev = [[int(x) for x in row] for row in J[c, j - 1]]
            assert [[int(x) for x in row] for row in J[c, j]] == matmul(prev, prev, m)


def test_substreams_by_big_integer_powers():
    seed = [777] * 6
    idx = [0, 1, 5, 1000, 2**33 + 17, 2**63 + 3]
    st = O.rng_init(np.array(seed, dtype=np.uint32), idx)
    p1 = [[int(x) for x in r] for r in GOLD["A1p76"]]
    p2 = [[int(x) for x in r] for r in GOLD["A2p76"]]
    for k, i in enumerate(idx):
        v1 = matvec(matpow(p1, i, M1), seed[:3], M1)
        v2 = matvec(matpow(p2, i, M2), seed[3:], M2)
        assert list(st[k]) == v1 + v2


def test_recurrence_matches_transition_matrices():
    st = O.rng_init(np.full(6, 12345, dtype=np.uint32), [3])
    s = [int(x) for x in st[0]]
    _, after = O.rng_draw(st, 37)
    v1 = matvec(matpow(A1, 37, M1), s[:3], M1)
    v2 = matvec(matpow(A2, 37, M2), s[3:], M2)
    assert list(after[0]) == v1 + v2


def test_output_formula_and_range():
    st = O.rng_init(np.full(6, 12345, dtype=np.uint32), np.arange(64, dtype=np.uint64))
    s = [[int(x) for x in row] for row in st]
    d, _ = O.rng_draw(st, 50)
    norm = 1.0 / (M1 + 1.0)
    for t in range(8):
        v = list(s[t])
        for j in range(50):
            p1 = (1403580 * v[1] - 810728 * v[0]) % M1
            p2 = (527612 * v[5] - 1370589 * v[3]) % M2
            v = [v[1], v[2], p1, v[4], v[5], p2]
            u = (p1 - p2) * norm if p1 > p2 else (p1 - p2 + M1) * norm
            assert d[t, j] == u
    assert d.min() > 0 and d.max() < 1
    # first two moments of 3200 draws
    assert abs(d.mean() - 0.5) < 0.02 and abs(d.var() - 1 / 12) < 0.01


def test_distinct_substreams_do_not_overlap():
    st = O.rng_init(np.full(6, 12345, dtype=np.uint32), np.arange(200, dtype=np.uint64))
    d, _ = O.rng_draw(st, 20)
    assert len(np.unique(d.ravel())) == d.size
