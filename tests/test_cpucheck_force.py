"""CPU check of the force kernel's per-site arithmetic (force_site.h compiled
for the host, looped over sites and directions) against the oracle's
lie_derivative (dirac.py:525-565): the closed-form generator sums, the spin
combinations, the adjoint commutator form, weights, symmetrisation,
accumulation and the boundary factor, for every (ndim, Nc, representation)
the kernel builds."""
import ctypes
import os

import numpy as np
import pytest

from oracle import force as ofo
from oracle import wilson as ow
from tfleonardym_amd import fields

LIB = os.path.join(os.path.dirname(__file__), "..", "tfleonardym_amd", "liblym_cpucheck.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from tfleonardym_amd import build
        build.build_cpu_check()
    lib = ctypes.CDLL(LIB)
    lib.lymcheck_lie_derivative.restype = ctypes.c_int
    return lib


def _ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data_as(ctypes.c_void_p) for a in arrs])


def run(lib, geometry, nc, adj, herm, kappa, bc, links, pairs, weights, sym, out=None):
    ndim = len(geometry)
    dims = (ctypes.c_int * 4)(*(list(geometry) + [1] * (4 - ndim)))
    V = links.shape[-1]
    accumulate = 0 if out is None else 1
    out = np.zeros((ndim, nc, nc, V), dtype=np.complex128) if out is None else np.array(out, copy=True)
    xs = [np.ascontiguousarray(p[0]) for p in pairs]
    ys = [np.ascontiguousarray(p[1]) for p in pairs]
    w = (ctypes.c_double * max(1, len(weights)))(*weights)
    links = np.ascontiguousarray(links)
    rc = lib.lymcheck_lie_derivative(ndim, nc, adj, herm, dims, ctypes.c_double(kappa), ctypes.c_double(bc.real),
                                     ctypes.c_double(bc.imag), links.ctypes.data_as(ctypes.c_void_p), None,
                                     len(pairs), _ptrs(xs), _ptrs(ys), w, sym, accumulate,
                                     out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    return out


CASES = []
for geometry in [(4, 6), (3, 4), (2, 4, 2, 4), (4, 2, 3, 4)]:
    for nc in (2, 3):
        for adj in (0, 1):
            CASES.append((geometry, nc, adj))


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("geometry,nc,adj", CASES)
@pytest.mark.parametrize("herm", [0, 1])
@pytest.mark.parametrize("bc", [None, -1.0, np.exp(0.3j)])
def test_force_site_arithmetic_matches_oracle(lib, geometry, nc, adj, herm, bc):
    D = len(geometry)
    ns = 2 if D == 2 else 4
    rep = "adjoint" if adj else "fundamental"
    U = fields.gauge_field(geometry, nc, seed=17, kind="haar")
    kappa = 0.1234
    wo = ow.WilsonOracle(U, geometry, kappa, hermitian=bool(herm), representation=rep, bc_factor=bc)
    rng = np.random.default_rng(6)
    V = U.shape[-1]
    shp = (ns, nc, V) if rep == "fundamental" else (ns, nc, nc, V)
    X = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    Y = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    b = 1.0 + 0j if bc is None else complex(bc)
    # single, unsymmetrised, unit weight: exactly lie_derivative(X, Y)
    got = run(lib, geometry, nc, adj, herm, kappa, b, U, [(X, Y)], [1.0], 0)
    assert rel(got, ofo.lie_derivative(wo, X, Y)) < 1e-13
    # two weighted symmetrised pairs, accumulated onto a prior field
    Z = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    prior = rng.standard_normal(got.shape) + 1j * rng.standard_normal(got.shape)
    got = run(lib, geometry, nc, adj, herm, kappa, b, U, [(X, Y), (Z, X)], [0.7, -1.3], 1, out=prior)
    want = prior + 0.7 * (ofo.lie_derivative(wo, X, Y) + ofo.lie_derivative(wo, Y, X)) \
        - 1.3 * (ofo.lie_derivative(wo, Z, X) + ofo.lie_derivative(wo, X, Z))
    assert rel(got, want) < 1e-13


def test_force_no_pairs_zeroes(lib):
    geometry = (4, 6)
    U = fields.gauge_field(geometry, 2, seed=1, kind="haar")
    got = run(lib, geometry, 2, 0, 1, 0.1, 1.0 + 0j, U, [], [], 1)
    assert not got.any()
