"""Pins the oracle's B-spline bases and de Boor evaluation (splines.hpp:39-110)
against exact rational Cox-de Boor recursion, the closed-form cubic pieces and
partition of unity."""
from fractions import Fraction

import numpy as np
import pytest

from tests import oracle_lib as O

TS = [0.0, 0.1, 0.25, 0.37, 0.5, 0.731, 0.99]


def cardinal(k: int, s: Fraction) -> Fraction:
    """Cardinal B-spline M_k(s) on [0, k) by the Cox-de Boor recursion (exact)."""
    if k == 1:
        return Fraction(1) if 0 <= s < 1 else Fraction(0)
    return (s * cardinal(k - 1, s) + (k - s) * cardinal(k - 1, s - 1)) / (k - 1)


def exact_basis(order: int, deriv: int, t: float):
    """Basis i on the cell [0,1) is M_order(t + order - 1 - i); d/ds M_k = M_{k-1}(s) - M_{k-1}(s-1)."""
    tf = Fraction(t)
    out = []
    for i in range(order):
        s = tf + order - 1 - i
        if deriv == 0:
            out.append(cardinal(order, s))
        elif deriv == 1:
            out.append(cardinal(order - 1, s) - cardinal(order - 1, s - 1))
        else:
            out.append(cardinal(order - 2, s) - 2 * cardinal(order - 2, s - 1) + cardinal(order - 2, s - 2))
    return np.array([float(v) for v in out])


@pytest.mark.parametrize("order", [3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("deriv", [0, 1, 2])
def test_basis_matches_exact_recursion(order, deriv):
    for t in TS:
        got = O.basis(order, deriv, t)
        ref = exact_basis(order, deriv, t)
        assert np.max(np.abs(got - ref)) <= 4e-15 * max(1.0, np.max(np.abs(ref))), (order, deriv, t, got, ref)


def test_cubic_closed_form():
    for t in TS:
        a = 1 - t
        ref = [a ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6, (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, t ** 3 / 6]
        dref = [-a ** 2 / 2, (3 * t ** 2 - 4 * t) / 2, (-3 * t ** 2 + 2 * t + 1) / 2, t ** 2 / 2]
        assert np.allclose(O.basis(4, 0, t), ref, rtol=0, atol=2e-16)
        assert np.allclose(O.basis(4, 1, t), dref, rtol=0, atol=4e-16)
    # collocation weights of fields.hpp:70-75
    assert np.array_equal(O.basis(4, 0, 0.0), np.array([1 / 6, 4 / 6, 1 / 6, 0.0]))


@pytest.mark.parametrize("order", [3, 4, 5, 6, 7, 8])
def test_partition_of_unity(order):
    for t in TS:
        assert abs(O.basis(order, 0, t).sum() - 1.0) < 4e-16
        assert abs(O.basis(order, 1, t).sum()) < 4e-15


@pytest.mark.parametrize("order", [3, 4, 5, 6, 8])
@pytest.mark.parametrize("deriv", [0, 1, 2])
def test_deboor_equals_basis_sum(order, deriv):
    rng = np.random.default_rng(order * 10 + deriv)
    for t in TS:
        c = rng.standard_normal(order)
        got = O.deboor(order, deriv, t, c)
        ref = float(np.dot(c, exact_basis(order, deriv, t)))
        assert abs(got - ref) <= 2e-14 * np.abs(c).sum(), (order, deriv, t)


def test_deboor_derivative_is_finite_difference():
    rng = np.random.default_rng(7)
    c = rng.standard_normal(4)
    h = 1e-6
    for t in [0.2, 0.5, 0.8]:
        fd = (O.deboor(4, 0, t + h, c) - O.deboor(4, 0, t - h, c)) / (2 * h)
        assert abs(O.deboor(4, 1, t, c) - fd) < 1e-8
