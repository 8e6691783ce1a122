"""SU(2)/SU(3) generators (lie_generators.py:3-37).

Fundamental: T^a = sigma^a/2 and lambda^a/2.  Adjoint: (T_adj^i)_{jk} =
-2 tr([T^i, T^j] T^k) from the structure constants.  Returned as NumPy
complex128 arrays, as in the reference.
"""
import numpy as np

from .fields import fundamental_generators

_cache = {}


def generators(nc, representation="fundamental"):
    key = (nc, representation)
    if key in _cache:
        return _cache[key]
    if representation == "fundamental":
        g = fundamental_generators(nc)
    elif representation == "adjoint":
        t = fundamental_generators(nc)
        n = nc * nc - 1
        g = np.zeros((n, n, n), dtype=np.complex128)
        for i in range(n):
            for j in range(n):
                comm = t[i] @ t[j] - t[j] @ t[i]
                for k in range(n):
                    g[i, j, k] = -2 * np.trace(comm @ t[k])
    else:
        raise KeyError(representation)
    _cache[key] = g
    return g
