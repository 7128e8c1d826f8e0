"""Pins the project double exp and the reaction probability against
correctly rounded values from Python's decimal module."""
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle

getcontext().prec = 50


def exact_exp(x: float) -> float:
    return float(Decimal(x).exp())


def ulps(a: float, b: float) -> int:
    ia = np.array([a]).view(np.int64)[0]
    ib = np.array([b]).view(np.int64)[0]
    return abs(int(ia) - int(ib))


def test_exp_within_one_ulp_and_almost_always_correctly_rounded():
    rng = np.random.default_rng(1)
    xs = np.concatenate([-rng.uniform(0, 2, 3000), -(10.0 ** rng.uniform(-9, 2.85, 3000)),
                         rng.uniform(-700, 700, 2000)])
    worst, exact = 0, 0
    for x in xs:
        got = oracle.exp(float(x))
        want = exact_exp(float(x))
        d = ulps(got, want)
        worst = max(worst, d)
        exact += d == 0
    assert worst <= 1
    assert exact / xs.size > 0.98


def test_exp_special_values():
    assert oracle.exp(0.0) == 1.0
    assert oracle.exp(-0.0) == 1.0
    assert oracle.exp(-746.0) == 0.0
    assert oracle.exp(710.0) == float("inf")
    assert np.isnan(oracle.exp(float("nan")))
    # subnormal range
    assert 0.0 < oracle.exp(-740.0) < 1e-320


@pytest.mark.parametrize("a", [1e-7, 2.534e-3, 5e-6, 0.0125, 0.3, 1.0, 7.5, 40.0])
def test_reaction_probability_closed_form(a):
    a32 = float(np.float32(a))
    want = float(np.float32(1.0 - exact_exp(-a32)))
    assert oracle.reaction_probability(a32) == want


def test_reaction_probability_monotone_and_bounded():
    a = np.float32(10.0) ** np.linspace(-8, 1.5, 2000, dtype=np.float32)
    p = np.array([oracle.reaction_probability(float(v)) for v in a], dtype=np.float32)
    assert np.all(np.diff(p) >= 0)
    assert p[0] > 0 and p[-1] <= 1.0
    # small a: 1 - exp(-a) ~ a
    assert abs(p[0] / a[0] - 1) < 1e-6
