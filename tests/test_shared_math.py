"""The FP64 functions shared by the oracle and the device build, pinned
against 60-digit values.

sin and cos (Cody-Waite reduction by pi/2 in three parts, Taylor
polynomials):
  - the split of pi/2 sums to pi/2 within 1e-36 and its first two parts hold
    33 bits (k P1, k P2 exact for |k| < 2^20); 2/pi is correctly rounded;
  - sin and cos are within 2 ulp of the true values (1 ulp = spacing at the
    result) over |x| <= 1e4 and at multiples of pi/4 up to 1e4 turns, and
    exactly odd / even;
  - sin^2 + cos^2 = 1 within 4e-16.

log, exp, log1p, expm1 (2 atanh series after reduction by powers of two;
Cody-Waite reduction by ln 2 and a Taylor series; Kahan's log1p correction):
  - within 2 ulp of the true values over their working ranges, and exact at
    the special points (log 1 = 0, exp 0 = 1, log 2^k = k ln 2 to 1 ulp);
  - pow(x, y) = exp(y log x) within 2 + |y log x| ulp (its only uses are
    cube roots), log10 within 3 ulp (spectrum bins only);
  - the edges: log(0) = -inf, log(<0) = NaN, exp overflow to inf and
    underflow to 0, exp(-745.1) a subnormal.
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


def _fn(which, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    yy = None if y is None else abi.ptr(np.ascontiguousarray(y, dtype=np.float64))
    O.check(O.lib().gxref_mathfn(which, abi.ptr(x), yy, abi.ptr(out), len(x)))
    return out


def _max_ulp(got, want):
    w = np.array([float(v) for v in want])
    sp = np.spacing(np.abs(w))
    return max(float(abs(mpmath.mpf(g) - v)) / s for g, v, s in zip(got, want, sp))


def test_log_exp_family_against_60_digits():
    mpmath.mp.dps = 60
    rs = np.random.default_rng(8)
    x = np.concatenate([np.exp(rs.uniform(-700, 700, 3000)), rs.uniform(0.5, 2.0, 2000), 1 + rs.uniform(-1e-6, 1e-6, 200),
                        2.0 ** np.arange(-1000, 1000, 37)])
    assert _max_ulp(_fn(0, x), [mpmath.log(mpmath.mpf(v)) for v in x]) <= 2.0
    xe = np.concatenate([rs.uniform(-745, 709.7, 3000), rs.uniform(-1, 1, 2000), rs.uniform(-1e-8, 1e-8, 200)])
    ok = np.abs(xe) < 708  # normal results
    assert _max_ulp(_fn(1, xe[ok]), [mpmath.exp(mpmath.mpf(v)) for v in xe[ok]]) <= 2.0
    x1 = np.concatenate([rs.uniform(-0.999, 10, 3000), np.exp(rs.uniform(np.log(1e-300), np.log(1e-2), 1000)),
                         -np.exp(rs.uniform(np.log(1e-300), np.log(0.5), 1000))])
    assert _max_ulp(_fn(2, x1), [mpmath.log1p(mpmath.mpf(v)) for v in x1]) <= 2.0
    xm = np.concatenate([rs.uniform(-30, 30, 2000), np.exp(rs.uniform(np.log(1e-300), np.log(0.3), 1000)),
                         -np.exp(rs.uniform(np.log(1e-300), np.log(0.3), 1000))])
    assert _max_ulp(_fn(3, xm), [mpmath.expm1(mpmath.mpf(v)) for v in xm]) <= 2.0
    xp = np.concatenate([np.arange(1.0, 101.0), rs.uniform(0, 1, 1000)])
    yp = np.full(len(xp), 1.0 / 3.0)
    got = _fn(4, xp, yp)
    want = [mpmath.mpf(v) ** mpmath.mpf(1.0 / 3.0) for v in xp]
    bound = 2.0 + np.abs(np.log(xp) / 3.0) / 1.0 + 1.0
    for g, w, b in zip(got, want, bound):
        assert float(abs(mpmath.mpf(g) - w)) <= b * np.spacing(float(w))
    assert _max_ulp(_fn(5, x), [mpmath.log10(mpmath.mpf(v)) for v in x]) <= 3.0
    assert _fn(0, [1.0])[0] == 0.0 and _fn(1, [0.0])[0] == 1.0
    e = _fn(0, [0.0, -1.0, np.inf])
    assert e[0] == -np.inf and np.isnan(e[1]) and e[2] == np.inf
    f = _fn(1, [710.0, -746.0, -745.1, np.nan])
    assert f[0] == np.inf and f[1] == 0.0 and 0.0 < f[2] < 1e-320 and np.isnan(f[3])
