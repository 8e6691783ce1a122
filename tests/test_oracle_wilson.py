"""Pins the oracle's Wilson operator (dirac.py:253-523, 649-662).

Checks are chosen so that a dropped term, a wrong sign, a wrong neighbour or a
transposed link fails at least one: explicit dense assembly by loops, the
free-field closed form, gauge covariance, gamma5-hermiticity, projected ==
simple and adjoint == fundamental-with-adjoint-links (tests.py:51-94)."""
import numpy as np
import pytest
from scipy.linalg import lu

from oracle import lattice as olat
from oracle import wilson as ow
from tfleonardym_amd import fields


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("ndim", [2, 4])
def test_gamma_algebra(ndim):
    g, g5 = ow.gammas(ndim)
    ns = g.shape[1]
    for mu in range(ndim):
        assert np.allclose(g[mu], g[mu].conj().T)
        assert np.allclose(g[mu] @ g5 + g5 @ g[mu], 0)
        for nu in range(ndim):
            ac = g[mu] @ g[nu] + g[nu] @ g[mu]
            assert np.allclose(ac, 2 * (mu == nu) * np.eye(ns))
    if ndim == 4:
        assert np.allclose(g5, g[0] @ g[1] @ g[2] @ g[3])
    else:
        assert np.allclose(g5, -1j * g[0] @ g[1])


@pytest.mark.parametrize("ndim", [2, 4])
@pytest.mark.parametrize("herm", [False, True])
def test_lu_reconstruction_exact(ndim, herm):
    """The LU back-projection reproduces the full projector for every mu/sign."""
    g, g5 = ow.gammas(ndim)
    ns = g.shape[1]
    h = ns // 2
    for mu in range(ndim):
        for sgn in (-1, 1):
            P = np.eye(ns) + sgn * g[mu]
            if herm:
                P = g5 @ P
            pl, _ = lu(P, permute_l=True)
            recon = np.zeros_like(P)
            recon[:h] = P[:h]
            recon[h:] = pl[h:, :h] @ P[:h]
            assert np.allclose(recon, P)


CASES = [((4, 4), 2), ((2, 2, 2, 2), 3), ((3, 4), 3), ((2, 2, 2, 4), 2)]


@pytest.mark.parametrize("geometry,nc", CASES)
@pytest.mark.parametrize("herm", [False, True])
@pytest.mark.parametrize("bc", [None, -1.0])
def test_dense_assembly(geometry, nc, herm, bc):
    D = len(geometry)
    ns = 2 if D == 2 else 4
    U = fields.gauge_field(geometry, nc, seed=11, kind="haar")
    op = ow.WilsonOracle(U, geometry, kappa=0.137, hermitian=herm, representation="fundamental", bc_factor=bc)
    M = ow.dense_matrix_fundamental(op.links, geometry, 0.137, herm)
    psi = fields.gaussian_spinor(geometry, ns, nc, seed=3)
    want = (M @ psi.reshape(-1)).reshape(psi.shape)
    assert rel(op.apply(psi), want) < 1e-14
    op.mode = "simple"
    assert rel(op.apply(psi), want) < 1e-14
    if herm:
        assert np.allclose(M, M.conj().T)  # Q is Hermitian


@pytest.mark.parametrize("ndim,geometry", [(2, (8, 6)), (4, (4, 2, 2, 6))])
@pytest.mark.parametrize("anti", [False, True])
def test_free_field_plane_wave(ndim, geometry, anti):
    nc = 2
    U = fields.gauge_field(geometry, nc, seed=0, kind="unit")
    kappa = 0.11
    op = ow.WilsonOracle(U, geometry, kappa, hermitian=False, bc_factor=-1.0 if anti else None)
    c = olat.coordinates(geometry)
    n = np.array([1, 2, 1, 3][:ndim])
    p = 2 * np.pi * n / np.array(geometry, dtype=float)
    if anti:
        p[-1] = (2 * n[-1] + 1) * np.pi / geometry[-1]
    phase = np.exp(1j * np.einsum("m,ms->s", p, c))
    ns = 2 if ndim == 2 else 4
    rng = np.random.default_rng(5)
    u = rng.standard_normal((ns, nc)) + 1j * rng.standard_normal((ns, nc))
    psi = u[:, :, None] * phase[None, None, :]
    M = ow.free_field_eigenvalue(p, kappa, ndim)
    want = np.einsum("st,tc->sc", M, u)[:, :, None] * phase[None, None, :]
    assert rel(op.apply(psi), want) < 1e-13


@pytest.mark.parametrize("geometry,nc", [((4, 6), 2), ((2, 4, 2, 4), 3)])
def test_gauge_covariance(geometry, nc):
    D = len(geometry)
    ns = 2 if D == 2 else 4
    U = fields.gauge_field(geometry, nc, seed=1, kind="haar")
    V = fields.volume(geometry)
    G = fields.haar_sun(np.random.default_rng(9), V, nc)  # (V, nc, nc)
    sup, _ = olat.lookup_tables(geometry)
    UG = np.empty_like(U)
    for mu in range(D):
        Um = U[mu].transpose(2, 0, 1)
        UG[mu] = np.einsum("sij,sjk,slk->sil", G, Um, np.conj(G[sup[mu]])).transpose(1, 2, 0)
    psi = fields.gaussian_spinor(geometry, ns, nc, seed=2)
    Gpsi = np.einsum("sij,mjs->mis", G, psi)
    op = ow.WilsonOracle(U, geometry, 0.12, hermitian=True, bc_factor=-1.0)
    opG = ow.WilsonOracle(UG, geometry, 0.12, hermitian=True, bc_factor=-1.0)
    lhs = opG.apply(Gpsi)
    rhs = np.einsum("sij,mjs->mis", G, op.apply(psi))
    assert rel(lhs, rhs) < 1e-13


@pytest.mark.parametrize("rep", ["fundamental", "adjoint"])
def test_gamma5_hermiticity(rep):
    geometry = (8, 8)
    nc = 2
    U = fields.gauge_field(geometry, nc, seed=4, kind="haar")
    op = ow.WilsonOracle(U, geometry, 0.1, hermitian=True, representation=rep, bc_factor=-1.0)
    v1 = fields.gaussian_spinor(geometry, 2, nc, 7, representation=rep)
    v2 = fields.gaussian_spinor(geometry, 2, nc, 8, representation=rep)
    a = np.vdot(v2, op.apply(v1))
    b = np.vdot(op.apply(v2), v1)
    assert abs(a - b) < 1e-12 * abs(a)


@pytest.mark.parametrize("geometry,nc", [((8, 8), 2), ((2, 2, 2, 4), 3), ((4, 4), 3)])
@pytest.mark.parametrize("herm", [False, True])
def test_adjoint_matches_fundamental_with_adjoint_links(geometry, nc, herm):
    """tests.py:83-90: simple == projected, and the matrix form equals the
    fundamental operator with real links V_ab on x^a = 2 tr(T^a X)."""
    D = len(geometry)
    ns = 2 if D == 2 else 4
    U = fields.gauge_field(geometry, nc, seed=21, kind="haar")
    X = fields.gaussian_spinor(geometry, ns, nc, 3, representation="adjoint")
    adj = ow.WilsonOracle(U, geometry, 0.1, hermitian=herm, representation="adjoint", bc_factor=-1.0)
    t_proj = adj.apply(X)
    adj.mode = "simple"
    t_simple = adj.apply(X)
    assert rel(t_proj, t_simple) < 1e-14
    Vl = ow.adjoint_links(U)
    assert np.abs(Vl.imag).max() < 1e-13
    fund = ow.WilsonOracle(Vl, geometry, 0.1, hermitian=herm, representation="fundamental", bc_factor=-1.0)
    t3 = fund.apply(ow.to_adjoint_vector(X))
    assert rel(t3, ow.to_adjoint_vector(t_proj)) < 1e-13
    # and against the dense matrix built from the adjoint links
    M = ow.dense_matrix_fundamental(fund.links, geometry, 0.1, herm)
    x = ow.to_adjoint_vector(X)
    assert rel(t3, (M @ x.reshape(-1)).reshape(x.shape)) < 1e-13


def test_adjoint_trace_part_hops_with_unit_links():
    """A trace component X = c(x) 1 is invariant under U X U^dag, so it sees
    the free operator (with the boundary sign once)."""
    geometry = (4, 6)
    nc = 2
    U = fields.gauge_field(geometry, nc, seed=2, kind="haar")
    V = fields.volume(geometry)
    rng = np.random.default_rng(0)
    c = rng.standard_normal((2, V)) + 1j * rng.standard_normal((2, V))
    X = np.einsum("ms,ij->mijs", c, np.eye(nc))
    adj = ow.WilsonOracle(U, geometry, 0.13, hermitian=True, representation="adjoint", bc_factor=-1.0)
    free = ow.WilsonOracle(fields.gauge_field(geometry, 1, 0, kind="unit"), geometry, 0.13,
                           hermitian=True, bc_factor=-1.0)
    want = np.einsum("ms,ij->mijs", free.apply(c[:, None, :])[:, 0, :], np.eye(nc))
    assert rel(adj.apply(X), want) < 1e-13


def test_square_operator_positive_definite():
    geometry = (4, 4)
    U = fields.gauge_field(geometry, 2, seed=5, kind="haar")
    op = ow.WilsonOracle(U, geometry, 0.15, hermitian=True, bc_factor=-1.0)
    M = ow.dense_matrix_fundamental(op.links, geometry, 0.15, True)
    A = M @ M
    assert np.allclose(A, A.conj().T)
    assert np.linalg.eigvalsh(A).min() > 0
    psi = fields.gaussian_spinor(geometry, 2, 2, seed=1)
    want = (A @ psi.reshape(-1)).reshape(psi.shape) + 0.3 * psi
    assert rel(op.apply_square(psi, 0.3), want) < 1e-13
