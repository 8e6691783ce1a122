"""Oracle: fermion force (test infrastructure only).

NumPy complex128 restatement of

* ``lie_generators.generators(nc, "adjoint")``  (lie_generators.py:20-37)
* ``DiracWilsonOperator.lie_derivative``        (dirac.py:525-565)
* ``fermion_action.n_flavor``                   (fermion_action.py:7-93):
  pseudofermions, energy, number of flavours, force

with the multishift solves done by textbook CG per pole (the reference's
MR multishift converges to the same solutions).

Conventions (from the reference): ``lie_derivative(X, Y)`` returns a
(D, Nc, Nc, V) field ``kappa * sum_k T^k_fund l_k`` with

    l_k = X^dag P+ (T^k_rep U_mu) Y(x+mu)  -  Y(x+mu)^dag P- (U_mu^dag T^k_rep) X

P+ = [g5] (1 - g_mu), P- = [g5] (1 + g_mu), U = ``fermion_gauge_field`` after
the boundary provider (the adjoint links V_ab for the adjoint
representation, whose spinors are first turned into x^a = 2 tr(T^a X)).

Pinning (tests/test_oracle_force.py): central finite differences of
X^dag Q(U_eps) Y + Y^dag Q(U_eps) X under U_mu(x0) -> exp(i eps T^a) U_mu(x0),
with Q the oracle Wilson operator (itself pinned by dense assembly), equal
i kappa (l_a(X,Y) + l_a(Y,X)) -- half that for the adjoint matrix form, whose
inner product is half the adjoint-vector one; and finite differences of the
n_flavor energy against the force it returns.
"""
import numpy as np

from . import lattice as olat
from . import solvers as osol
from . import wilson as ow


def adjoint_generators(nc):
    """(T^i_adj)_{jk} = -2 tr([T^i, T^j] T^k) = -i f_ijk (lie_generators.py:20-37)."""
    T = ow.generators(nc)
    n = nc * nc - 1
    G = np.zeros((n, n, n), dtype=np.complex128)
    for i in range(n):
        for j in range(n):
            C = T[i] @ T[j] - T[j] @ T[i]
            for k in range(n):
                G[i, j, k] = -2.0 * np.trace(C @ T[k])
    return G


def fermion_gauge_field(wo):
    """U after the boundary provider, or V_ab = 2 tr(T^a U T^b U^dag) of the
    original links with the provider applied to V (dirac.py:356-370)."""
    if wo.representation == "fundamental":
        return wo.links
    V = ow.adjoint_links(wo.links_orig)
    if wo.bc_factor is None or wo.bc_factor == 1:
        return V
    return olat.apply_antiperiodic(V, wo.geometry, wo.bc_factor)


def lie_derivative(wo, X, Y):
    """DiracWilsonOperator.lie_derivative (dirac.py:525-565)."""
    X = np.asarray(X, dtype=np.complex128)
    Y = np.asarray(Y, dtype=np.complex128)
    nc = wo.links_orig.shape[1]
    if X.ndim == 4:
        X = ow.to_adjoint_vector(X)
        Y = ow.to_adjoint_vector(Y)
    D = wo.links_orig.shape[0]
    g, g5 = ow.gammas(D)
    eye = np.eye(g.shape[1])
    Tf = ow.generators(nc)
    Trep = adjoint_generators(nc) if wo.representation == "adjoint" else Tf
    links = fermion_gauge_field(wo)
    out = []
    for mu in range(D):
        plus = eye - g[mu]
        minus = eye + g[mu]
        if wo.hermitian:
            plus = g5 @ plus
            minus = g5 @ minus
        Ysup = olat.translate(Y, mu, +1, wo.tables)
        d = np.einsum("mis,mn,kij,jls,nls,kut->uts", np.conj(X), plus, Trep, links[mu], Ysup, Tf,
                      optimize=True)
        d = d - np.einsum("mis,mn,jis,kjl,nls,kut->uts", np.conj(Ysup), minus, np.conj(links[mu]), Trep, X, Tf,
                          optimize=True)
        out.append(d)
    return wo.kappa * np.stack(out)


# ---------------------------------------------------------------------------
# n_flavor (fermion_action.py:7-93) with r(A) v = shift v + sum alpha_i x_i,
# (A + beta_i) x_i = v solved by CG (beta_i < 0 as (A - |beta_i|) would not be
# positive definite; the tables used here have beta_i >= 0)
# ---------------------------------------------------------------------------

def _square(wo):
    return lambda v: wo.apply(wo.apply(v))


def shifted_solutions(wo, v, betas, tol=1e-12):
    A = _square(wo)
    out = []
    for b in betas:
        x, _ = osol.cg(A, v, shift=float(b), tol=tol)
        out.append(x)
    return out


def rational_apply(wo, v, alphas, betas, shift, tol=1e-12):
    """RationalApproximation.apply (dirac.py:52-59)."""
    xs = shifted_solutions(wo, v, betas, tol)
    r = shift * np.asarray(v, dtype=np.complex128)
    for a, x in zip(alphas, xs):
        r = r + a * x
    return r


def energy(wo, phis, approximations, tol=1e-12):
    """Re sum_k <phi_k, R_k(A) phi_k> (fermion_action.py:39-44);
    approximations: list of (alphas, betas, shift)."""
    e = 0.0
    for phi, (al, be, sh) in zip(phis, approximations):
        e += np.vdot(phi, rational_apply(wo, phi, al, be, sh, tol)).real
    return e


def number_of_flavors(approximations):
    """2 log prod r_k(0.5) / log 2 (fermion_action.py:46-52)."""
    r = 1.0
    for al, be, sh in approximations:
        r *= sh + sum(a / (0.5 + b) for a, b in zip(al, be))
    return 2.0 * np.log(r) / np.log(2.0)


def force(wo, phis, approximations, tol=1e-12):
    """sum_k sum_i alpha_i [L(x_i, Q x_i) + L(Q x_i, x_i)] with
    (A + beta_i) x_i = phi_k (fermion_action.py:64-93)."""
    D = wo.links_orig.shape[0]
    nc = wo.links_orig.shape[1]
    V = wo.links_orig.shape[-1]
    F = np.zeros((D, nc, nc, V), dtype=np.complex128)
    for phi, (al, be, sh) in zip(phis, approximations):
        xs = shifted_solutions(wo, phi, be, tol)
        for a, x in zip(al, xs):
            y = wo.apply(x)
            F = F + a * (lie_derivative(wo, x, y) + lie_derivative(wo, y, x))
    return F


def perturbed_links(links, mu, site, T, eps):
    """U_mu(site) -> exp(i eps T) U_mu(site) (T Hermitian; exact via eigh)."""
    w, Q = np.linalg.eigh(T)
    R = (Q * np.exp(1j * eps * w)) @ Q.conj().T
    out = np.array(links, dtype=np.complex128, copy=True)
    out[mu, :, :, site] = R @ out[mu, :, :, site]
    return out
