"""The correctly rounded pow / x^1.5 / hypot / sin / cos of the plain path's device arithmetic
(phygas_amd/csrc/crmath.h), through the host compilation of the same source (no device calls).

Pinned against the mathematics: every result must be the double nearest the exact value, which mpmath
gives at 200 bits, on random arguments over the ranges the force and the coordinates use and on cases with
exact answers.  Then the agreement with the C library (glibc, the reference's libm) is counted: it must be
at least 99%, the share where glibc itself rounds to nearest."""
import ctypes as C
from fractions import Fraction

import mpmath as mp
import numpy as np
import pytest

from phygas_amd import hip

mp.mp.prec = 200
LIBM = C.CDLL("libm.so.6")
for _n, _k in (("pow", 2), ("hypot", 2), ("sin", 1), ("cos", 1)):
    getattr(LIBM, _n).restype = C.c_double
    getattr(LIBM, _n).argtypes = [C.c_double] * _k


def _nearest(v) -> float:
    sign, man, exp, _ = mp.mpf(v)._mpf_
    f = Fraction(int(man)) * Fraction(2) ** exp
    return float(-f if sign else f)   # Fraction -> float rounds to nearest


def _exact(fn, x, y):
    X, Y = mp.mpf(float(x)), mp.mpf(float(y))
    if fn == "pow":
        return _nearest(mp.power(X, Y))
    if fn == "pow15":
        return _nearest(mp.power(X, mp.mpf(3) / 2))
    if fn == "hypot":
        return _nearest(mp.sqrt(X * X + Y * Y))
    return _nearest(mp.sin(X) if fn == "sin" else mp.cos(X))


def _glibc(fn, x, y):
    if fn == "pow":
        return LIBM.pow(x, y)
    if fn == "pow15":
        return LIBM.pow(x, 1.5)
    if fn == "hypot":
        return LIBM.hypot(x, y)
    return LIBM.sin(x) if fn == "sin" else LIBM.cos(x)


def _args(fn, n, seed):
    """Arguments over the plain path's ranges: the halo's x = r / 12 kpc (r >= 1e-12) with gamma and gamma - 1,
    the bulge's r^2 + b^2, the rotation prior's (x, y), right ascensions and declinations in radians."""
    rng = np.random.default_rng(seed)
    if fn == "pow":
        return np.exp(rng.uniform(np.log(8e-14), np.log(9.0), n)), np.where(rng.random(n) < 0.5, 2.02, 1.02)
    if fn == "pow15":
        return np.exp(rng.uniform(np.log(0.15), np.log(1e4), n)), np.zeros(n)
    if fn == "hypot":
        return rng.uniform(-30, 30, n), rng.uniform(-30, 30, n)
    return rng.uniform(-7.0, 7.0, n), np.zeros(n)


FUNCS = ["pow", "pow15", "hypot", "sin", "cos"]


@pytest.mark.parametrize("fn", FUNCS)
def test_correctly_rounded_against_mpmath(fn):
    x, y = _args(fn, 4000, 11 + FUNCS.index(fn))
    got = hip.crmath(fn, x, y)
    want = np.array([_exact(fn, a, b) for a, b in zip(x, y)])
    assert np.array_equal(got, want), int(np.sum(got != want))


def test_exact_cases():
    assert hip.crmath("pow", [4.0, 9.0, 1.0, 2.0, 0.5], [1.5, 0.5, 2.02, 10.0, 3.0]).tolist() == [8.0, 3.0, 1.0, 1024.0, 0.125]
    assert hip.crmath("pow15", [4.0, 0.25, 16.0, 1.0]).tolist() == [8.0, 0.125, 64.0, 1.0]
    assert hip.crmath("hypot", [3.0, 5.0, 0.0, -8.0], [4.0, 12.0, 0.0, 6.0]).tolist() == [5.0, 13.0, 0.0, 10.0]
    assert hip.crmath("sin", [0.0, -0.0]).tolist() == [0.0, -0.0]
    assert hip.crmath("cos", [0.0]).tolist() == [1.0]
    # a multiple of pi/2 is never exact in binary: the results are the tiny exact values
    x = np.array([np.pi / 2, np.pi, 3 * np.pi / 2, 2 * np.pi])
    want_s = np.array([_exact("sin", a, 0.0) for a in x])
    want_c = np.array([_exact("cos", a, 0.0) for a in x])
    assert np.array_equal(hip.crmath("sin", x), want_s)
    assert np.array_equal(hip.crmath("cos", x), want_c)


@pytest.mark.parametrize("fn", FUNCS)
def test_agreement_with_the_c_library(fn):
    x, y = _args(fn, 20000, 101 + FUNCS.index(fn))
    got = hip.crmath(fn, x, y)
    ref = np.array([_glibc(fn, a, b) for a, b in zip(x, y)])
    diff = got != ref
    assert diff.mean() <= 0.01, diff.mean()
    # where they differ, it is by one unit in the last place
    ulp = np.spacing(np.abs(ref[diff]))
    assert np.all(np.abs(got[diff] - ref[diff]) <= ulp)
