"""Pins the oracle's Boys function (src/pbasis.f90:1-27 table, Taylor /
asymptotic evaluation of src/pfourindexdirect.f90:543-553) against what the
mathematics fixes: F_m(0) = 1/(2m+1), F_0(T) = sqrt(pi/T) erf(sqrt T)/2, the
downward recursion F_{m-1} = (2T F_m + e^-T)/(2m-1), and the incomplete-gamma
closed form for every m used up to (dd|dd) classes."""
import math

import numpy as np
import pytest

import oracle
from tests.md_reference import boys as boys_exact


def test_table_grid_values_are_boys():
    t = oracle.boys_table()
    assert t.shape == (1701, 28)
    for m in range(28):
        assert abs(t[0, m] - 1.0 / (2 * m + 1)) < 1e-15
    for i in [1, 7, 20, 133, 700, 1700]:
        T = i / 20.0
        for m in [0, 1, 4, 8, 13]:
            ref = boys_exact(m, T)
            assert abs(t[i, m] - ref) <= 1e-14 * max(ref, 1e-300) + 1e-300, (i, m)


@pytest.mark.parametrize("m", range(0, 9))
def test_boys_against_incomplete_gamma(m):
    rng = np.random.default_rng(m)
    for T in np.concatenate([rng.uniform(0, 2, 40), rng.uniform(2, 40, 40), rng.uniform(40, 120, 20)]):
        ref = boys_exact(m, T)
        got = oracle.boys(T, m)
        assert abs(got - ref) <= 2e-14 * ref, (m, T, got, ref)


def test_f0_erf_closed_form():
    for T in [1e-3, 0.3, 1.0, 7.7, 25.0, 32.9, 33.1, 60.0]:
        ref = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
        assert abs(oracle.boys(T, 0) - ref) <= 2e-14 * ref


def test_downward_recursion_identity():
    for T in [0.01, 0.5, 3.3, 12.0, 29.0]:
        for m in range(1, 9):
            lhs = oracle.boys(T, m - 1)
            rhs = (2 * T * oracle.boys(T, m) + math.exp(-T)) / (2 * m - 1)
            assert abs(lhs - rhs) <= 1e-14 * lhs
