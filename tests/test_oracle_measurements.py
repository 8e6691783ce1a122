"""Pins the oracle's fermion measurements (fermion_measurements.py:41-115,
dirac.py:665-696): the propagator inverts Q up to g5, the per-slice sums
against an independent reshape of the site density, the free-field
correlator's reflection symmetry, and the condensate estimator against
tr(g5 Q^-1) from the dense operator."""
import numpy as np
import pytest

from oracle import measurements as om
from oracle import wilson as ow
from tfleonardym_amd import fields


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("geometry,nc,rep", [((4, 6), 2, "fundamental"), ((2, 2, 2, 4), 3, "fundamental"),
                                             ((4, 4), 2, "adjoint")])
def test_propagator_inverts_q(geometry, nc, rep):
    D = len(geometry)
    ns = 2 if D == 2 else 4
    U = fields.gauge_field(geometry, nc, seed=3, kind="near_identity")
    wo = ow.WilsonOracle(U, geometry, 0.12, hermitian=True, representation=rep, bc_factor=-1.0)
    src = fields.gaussian_spinor(geometry, ns, nc, seed=4, representation=rep)
    sol = om.propagator(wo, src)
    assert rel(wo.apply(ow.gamma5(sol, True)), src) < 1e-11


@pytest.mark.parametrize("geometry,rep", [((4, 6), "fundamental"), ((2, 3, 2, 4), "fundamental"),
                                          ((3, 4), "adjoint"), ((2, 2, 3, 2), "adjoint")])
def test_slice_sums_match_reshape(geometry, rep):
    D = len(geometry)
    ns = 2 if D == 2 else 4
    nc = 2
    V = fields.volume(geometry)
    rng = np.random.default_rng(1)
    shp = (ns, nc, V) if rep == "fundamental" else (ns, nc, nc, V)
    sols = [rng.standard_normal(shp) + 1j * rng.standard_normal(shp) for _ in range(3)]
    dens = sum(om.site_density(v, rep) for v in sols).reshape(geometry)
    for direction in range(D):
        axes = tuple(a for a in range(D) if a != direction)
        want = dens.sum(axis=axes)
        got = om.slice_sums(sols, geometry, direction, rep)
        assert np.allclose(got, want, rtol=1e-13, atol=0)
        assert np.allclose(got.imag, 0, atol=1e-12 * abs(got).max())
        if rep == "fundamental":
            assert abs(got.sum() - sum(np.vdot(v, v) for v in sols)) < 1e-12 * abs(got.sum())


def test_free_correlator_reflection_symmetry():
    """Unit links, periodic: the point-source correlator is symmetric under
    z -> L - z in every direction."""
    geometry = (6, 8)
    U = fields.gauge_field(geometry, 2, seed=0, kind="unit")
    wo = ow.WilsonOracle(U, geometry, 0.2, hermitian=True, representation="fundamental")
    for direction in range(2):
        c = om.pion_correlator(wo, direction)
        L = geometry[direction]
        for z in range(1, L):
            assert abs(c[z] - c[L - z]) < 1e-10 * abs(c[0])
        assert np.all(c.real > 0)


def test_condensate_estimator_is_trace_on_complete_sources():
    """mean over all unit sources e of <e, g5 Q^-1 e> / V = tr(g5 Q^-1) / (N V)."""
    geometry = (4, 4)
    nc = 2
    U = fields.gauge_field(geometry, nc, seed=9, kind="haar")
    wo = ow.WilsonOracle(U, geometry, 0.11, hermitian=True, representation="fundamental", bc_factor=-1.0)
    V = 16
    N = 2 * nc * V
    sources = []
    for k in range(N):
        e = np.zeros(N, dtype=np.complex128)
        e[k] = 1.0
        sources.append(e.reshape(2, nc, V))
    got = om.chiral_condensate(wo, sources)
    M = ow.dense_matrix_fundamental(wo.links, geometry, 0.11, True)
    G5 = np.kron(np.diag([1.0, -1.0]), np.eye(nc * V))
    want = np.trace(G5 @ np.linalg.inv(M)) / (N * V)
    assert abs(got - want) < 1e-11 * abs(want)
