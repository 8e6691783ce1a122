"""Rational tables: the values the reference documents for them.

Each table approximates a power of x; evaluating at a few points against
x**p pins that the alpha/beta split and the data are intact."""
import numpy as np
import pytest

from tfleonardym_amd import approximations as ap


@pytest.mark.parametrize("name,power,npoles", [
    ("heatbath_rational_approximation_coefficients", 1.0 / 16, 38),
    ("metropolis_rational_approximation_coefficients", -1.0 / 8, 17),
])
def test_tables_approximate_their_power(name, power, npoles):
    t = getattr(ap, name)
    a, b = ap.split(t)
    assert len(a) == len(b) == npoles
    for x in (1e-3, 0.05, 0.5, 2.0, 10.0):
        assert abs(ap.evaluate(t, x) / x ** power - 1) < 1e-3, x


def test_force_level1_is_coarse_inverse_eighth_root():
    t = ap.force_level1_rational_approximation_coefficients
    assert abs(ap.evaluate(t, 0.5) - 0.5 ** -0.125) < 1e-3


def test_heatbath_has_one_negative_pole_first():
    a, b = ap.split(ap.heatbath_rational_approximation_coefficients)
    assert b[0] < 0 and all(x > 0 for x in b[1:])
    # the base (last) pole is the smallest non-negative one
    assert b[-1] == min(b[1:])


def test_number_of_flavors_is_half():
    """fermion_action.py:46-52: 2 log(prod r_met(0.5)) / log 2 with 2 pseudofermions = 1/2."""
    r = ap.evaluate(ap.metropolis_rational_approximation_coefficients, 0.5)
    nf = 2 * np.log(r * r) / np.log(2.0)
    assert abs(nf - 0.5) < 1e-3


def test_square_root_polynomial_tables():
    for n in (40, 80):
        c = ap.get_square_root_polynomial_approximation_coefficients(n)
        assert len(c) == n + 1
    with pytest.raises(RuntimeError):
        ap.get_square_root_polynomial_approximation_coefficients(7)
