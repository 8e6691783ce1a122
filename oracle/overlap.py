"""Oracle: PolynomialApproximation and Overlap (test infrastructure only).

NumPy restatement of dirac.py:76-172: the polynomial
P(A) v = prod_roots scaling (A - root) v applied root by root, the overlap
operator O v = (1+m)/2 G v + (1-m)/2 Op P(Op^2) v with G = g5 (hermitian
form: g5 on the identity term; otherwise on the Op term), and its
lie_derivative (the reachable part, dirac.py:143-172).
"""
import numpy as np

from . import force as ofo
from . import wilson as ow


def poly_apply(wo, roots, scaling, v):
    """dirac.py:83-88 on A = Op^2."""
    out = np.array(v, dtype=np.complex128, copy=True)
    for r in roots:
        out = scaling * (wo.apply(wo.apply(out)) - r * out)
    return out


def overlap_apply(wo, roots, scaling, mass, hermitian, v):
    """dirac.py:120-141."""
    v = np.asarray(v, dtype=np.complex128)
    out = wo.apply(poly_apply(wo, roots, scaling, v))
    a = (1 + mass) / 2
    b = (1 - mass) / 2
    if hermitian:
        return a * ow.gamma5(v, True) + b * out
    return a * v + b * ow.gamma5(out, True)


def overlap_lie_derivative(wo, roots, scaling, mass, X, Y):
    """dirac.py:143-172: (1-m)/2 [L(X, P Y) + s sum_i (L(Op l_i, r_{i+1}) + L(l_i, Op r_{i+1}))]
    with l_0 = Op X, l_{i+1} = s (A - conj(root_i)) l_i and r_i the right
    products over the roots after i."""
    A = lambda v: wo.apply(wo.apply(v))
    factor = 0.5 - mass / 2.0
    right = [np.asarray(Y, dtype=np.complex128)]
    for r in reversed(list(roots)):
        right.append(scaling * (A(right[-1]) - r * right[-1]))
    right = list(reversed(right))
    left = [wo.apply(np.asarray(X, dtype=np.complex128))]
    for r in roots:
        left.append(scaling * (A(left[-1]) - np.conj(r) * left[-1]))
    F = factor * ofo.lie_derivative(wo, X, right[0])
    for i in range(len(roots)):
        F = F + (scaling * factor) * ofo.lie_derivative(wo, wo.apply(left[i]), right[i + 1])
        F = F + (scaling * factor) * ofo.lie_derivative(wo, left[i], wo.apply(right[i + 1]))
    return F
