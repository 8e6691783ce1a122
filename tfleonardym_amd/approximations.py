"""Rational and polynomial approximation tables (approximations.py:4-139).

The coefficient tables are data, loaded from ``data/rational_tables.json``
(extracted from the reference's literal lists by tests/golden/extract_tables.py).
Each rational table is the concatenation alpha || beta; ``split`` cuts it at
len//2 as the reference's driver does (TfLeonardYMNDim.py:310-332).
"""
import json
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "rational_tables.json")
with open(_DATA) as _f:
    _T = json.load(_f)

heatbath_rational_approximation_coefficients = np.array(_T["rational"]["heatbath_rational_approximation_coefficients"])
force_level1_rational_approximation_coefficients = np.array(
    _T["rational"]["force_level1_rational_approximation_coefficients"])
force_level2_rational_approximation_coefficients = np.array(
    _T["rational"]["force_level2_rational_approximation_coefficients"])
force_level3_rational_approximation_coefficients = np.array(
    _T["rational"]["force_level3_rational_approximation_coefficients"])
metropolis_rational_approximation_coefficients = np.array(
    _T["rational"]["metropolis_rational_approximation_coefficients"])


def split(table):
    """(alphas, betas) = table[:n//2], table[n//2:] as Python float lists."""
    n = len(table)
    return [float(a) for a in table[: n // 2]], [float(b) for b in table[n // 2:]]


def evaluate(table, x, shift=0.0):
    """shift + sum alpha_i / (x + beta_i) (dirac.py:61-67)."""
    a, b = split(table)
    return shift + sum(ai / (x + bi) for ai, bi in zip(a, b))


def get_square_root_polynomial_approximation_coefficients(n):
    """Scaling followed by the complex roots of the overlap polynomial
    (approximations.py:4-76); n in (40, 80)."""
    key = str(int(n))
    if key not in _T["square_root_polynomial"]:
        raise RuntimeError("Polynomial approximations", n, "not found")
    return np.array([c[0] + 1j * c[1] for c in _T["square_root_polynomial"][key]])
