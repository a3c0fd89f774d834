"""How far last-bit differences in the C library's transcendentals move the plain path's chi2 minima, measured
on the CPU with the oracle itself (liboracle_ulp.so: the same restatement with pow / sin / cos / hypot moved by
one unit in the last place in a chosen share of calls, a mistake planted in the oracle).

The tracks' state-transition matrix is a finite difference, A = (x_j - x_0) / 1e-5, so one ulp of an 8-kpc
position becomes ~1e-10 of A, and the chi2 minima carry it.  These tests fix the scale of that floor:
  * a libm whose pow misrounds half its arguments (the device library's, before the plain path took the
    correctly rounded pow of crmath.h) moves chi2 by ~1e-9 relative on the 20 x 400 golden inputs: the
    distribution measured on the GPU then (max 2.7e-9, 0.15% of pairs above 1e-9);
  * a pow that differs from glibc in 0.1% of calls (a correctly rounded pow: glibc's own misrounding share)
    stays far below 1e-9;
  * a one-ulp change in sin / cos moves start states, and those move chi2 by ~1e-9 too: so the plain path
    takes each object's nominal start state from the host's C library."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle.oracle_py import Oracle
from phygas_amd import io as pio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INP = os.path.join(ROOT, "tests", "golden", "inputs")
POW, SIN, COS, HYPOT = 1, 2, 4, 8


@pytest.fixture(scope="module")
def inputs():
    return pio.read_clusters(os.path.join(INP, "clusters400.txt")), pio.read_targets(os.path.join(INP, "stars20.txt"))


@pytest.fixture(scope="module")
def base(inputs):
    cl, st = inputs
    return Oracle().chi2_matrix(cl, st)


@pytest.fixture(scope="module")
def ulp():
    path = os.path.join(ROOT, "oracle", "liboracle_ulp.so")
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: run `make -C oracle`")
    o = Oracle(path)
    o.lib.orc_ulp_set.argtypes = [C.c_uint, C.c_double]
    o.lib.orc_ulp_set.restype = None
    yield o
    o.lib.orc_ulp_set(0, 0.0)


def _moved(ulp, inputs, base, mask, rate):
    cl, st = inputs
    ulp.lib.orc_ulp_set(mask, rate)
    c1, k1, s1 = ulp.chi2_matrix(cl, st)
    ulp.lib.orc_ulp_set(0, 0.0)
    c0, k0, s0 = base
    fin = np.isfinite(c0)
    assert np.array_equal(fin, np.isfinite(c1))
    return np.abs(c1[fin] - c0[fin]) / c0[fin]


def test_unmoved_build_is_the_oracle(ulp, inputs, base):
    rel = _moved(ulp, inputs, base, 0, 0.0)
    assert rel.max() == 0.0


def test_pow_misrounding_half_the_calls_gives_the_1e9_floor(ulp, inputs, base):
    rel = _moved(ulp, inputs, base, POW, 0.5)
    assert 1e-9 < rel.max() < 1e-8, rel.max()
    assert 1e-4 < np.mean(rel > 1e-9) < 1e-2
    assert np.mean(rel == 0) < 0.9


def test_pow_differing_in_a_thousandth_of_calls_stays_below_the_bar(ulp, inputs, base):
    rel = _moved(ulp, inputs, base, POW, 1e-3)
    assert rel.max() < 1e-9, rel.max()
    assert np.mean(rel == 0) > 0.95


def test_start_state_transcendentals_reach_the_floor_too(ulp, inputs, base):
    rel = _moved(ulp, inputs, base, SIN | COS | HYPOT, 0.5)
    assert rel.max() > 5e-10, rel.max()
