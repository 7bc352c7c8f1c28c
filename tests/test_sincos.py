"""The FP64 sine and cosine shared by the oracle and the device build
(Cody-Waite reduction by pi/2 in three parts, Taylor polynomials): pinned
against 60-digit values.
  - the split of pi/2 sums to pi/2 within 1e-36 and its first two parts hold
    33 bits (k P1, k P2 exact for |k| < 2^20); 2/pi is correctly rounded;
  - sin and cos are within 2 ulp of the true values (1 ulp = spacing at the
    result) over |x| <= 1e4 and at multiples of pi/4 up to 1e4 turns, and
    exactly odd / even;
  - sin^2 + cos^2 = 1 within 4e-16.
"""
import ctypes as C

import mpmath
import numpy as np

from geantexascalepilot_amd import abi
import oracle_harness as O


def _sincos(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c = np.zeros_like(x), np.zeros_like(x)
    O.check(O.lib().gxref_sincos(abi.ptr(x), abi.ptr(s), abi.ptr(c), len(x)))
    return s, c


def test_reduction_constants():
    mpmath.mp.dps = 60
    h = mpmath.pi / 2
    p1, p2, p3 = float.fromhex("0x1.921fb544p+0"), float.fromhex("0x1.0b4611a6p-34"), float.fromhex(
        "0x1.3198a2e037073p-69")
    assert abs(mpmath.mpf(p1) + mpmath.mpf(p2) + mpmath.mpf(p3) - h) < mpmath.mpf("1e-36")
    for p in (p1, p2):
        m, _ = np.frexp(p)
        assert float(m * 2.0 ** 33) == int(m * 2.0 ** 33)  # 33 significant bits
    assert float.fromhex("0x1.45f306dc9c883p-1") == float(2 / mpmath.pi)


def _ulp_err(got, want):
    want_f = np.array([float(w) for w in want])
    spacing = np.spacing(np.maximum(np.abs(want_f), 1e-300))
    return np.array([abs(mpmath.mpf(g) - w) for g, w in zip(got, want)], dtype=float) / spacing


def test_accuracy_against_60_digits():
    mpmath.mp.dps = 60
    rs = np.random.default_rng(5)
    x = np.concatenate([rs.uniform(-np.pi, np.pi, 3000), rs.uniform(-1e4, 1e4, 2000),
                        np.exp(rs.uniform(np.log(1e-12), np.log(1.0), 500)),
                        np.arange(-40000, 40001, 97) * (np.pi / 4), [0.0, 1e-300, -1e-300, np.pi / 2, np.pi]])
    s, c = _sincos(x)
    ws = [mpmath.sin(mpmath.mpf(v)) for v in x]
    wc = [mpmath.cos(mpmath.mpf(v)) for v in x]
    # away from the zeros of the function, where an absolute error of a few
    # 1e-16 is several ulp of a tiny result
    es, ec = _ulp_err(s, ws), _ulp_err(c, wc)
    big_s = np.abs(np.array([float(v) for v in ws])) > 1e-3
    big_c = np.abs(np.array([float(v) for v in wc])) > 1e-3
    assert es[big_s].max() <= 2.0, es[big_s].max()
    assert ec[big_c].max() <= 2.0, ec[big_c].max()
    # near zeros: absolute error within 2 ulp of the argument's reduction scale
    assert max(abs(float(a - b)) for a, b in zip(s, ws)) < 4e-16 * max(1.0, np.abs(x).max() / 4)
    s2, c2 = _sincos(-x)
    assert np.array_equal(s2, -s) and np.array_equal(c2, c)
    np.testing.assert_allclose(s * s + c * c, 1.0, atol=4e-16)
