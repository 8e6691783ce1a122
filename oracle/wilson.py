"""Oracle: Wilson-Dirac operator (test infrastructure only).

NumPy complex128 restatement of the reference ``DiracWilsonOperator``
(dirac.py:253-523, 649-662).  Conventions (all from the reference):

    D psi(x) = psi(x) + kappa * sum_mu [ (1 - g_mu) U_mu(x) psi(x+mu)
                                       + (1 + g_mu) U_mu(x-mu)^dag psi(x-mu) ]
    Q = g5 D  when hermitian                         (dirac.py:649-662)

Fundamental spinors are (Ns, Nc, V); adjoint spinors are (Ns, Nc, Nc, V)
matrices; links are (D, Nc, Nc, V).  ``links`` below always means the field
*after* the boundary-condition provider (``cfg`` / ``fermion_gauge_field``
of the reference) and ``links_orig`` the field before it.
"""
import numpy as np
from scipy.linalg import lu

from . import lattice as olat

# Gamma matrices transcribed from dirac.py:283-297 (chiral basis).
_G2 = np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]]], dtype=np.complex128)
_G5_2 = np.diag([1, -1]).astype(np.complex128)
_G4 = np.array([[[0, 0, 0, -1j], [0, 0, -1j, 0], [0, 1j, 0, 0], [1j, 0, 0, 0]],
                [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],
                [[0, 0, -1j, 0], [0, 0, 0, 1j], [1j, 0, 0, 0], [0, -1j, 0, 0]],
                [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]]], dtype=np.complex128)
_G5_4 = np.diag([1, 1, -1, -1]).astype(np.complex128)


def gammas(ndim):
    """(gamma[D], gamma5) for D in {2, 4}; RuntimeError otherwise (dirac.py:299-300)."""
    if ndim == 2:
        return _G2.copy(), _G5_2.copy()
    if ndim == 4:
        return _G4.copy(), _G5_4.copy()
    raise RuntimeError(f"Number of dimensions {ndim} not supported for fermions!")


# Lie-algebra generators T^a = sigma^a/2, lambda^a/2 (lie_generators.py:3-16).
def generators(nc):
    if nc == 2:
        g = np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]]])
    elif nc == 3:
        s3 = 1.0 / np.sqrt(3.0)
        g = np.array([[[0, 1, 0], [1, 0, 0], [0, 0, 0]],
                      [[0, 0, 1], [0, 0, 0], [1, 0, 0]],
                      [[0, 0, 0], [0, 0, 1], [0, 1, 0]],
                      [[0, -1j, 0], [1j, 0, 0], [0, 0, 0]],
                      [[0, 0, -1j], [0, 0, 0], [1j, 0, 0]],
                      [[0, 0, 0], [0, 0, -1j], [0, 1j, 0]],
                      [[1, 0, 0], [0, -1, 0], [0, 0, 0]],
                      [[s3, 0, 0], [0, s3, 0], [0, 0, -2 * s3]]])
    else:
        raise KeyError(nc)
    return (g / 2.0).astype(np.complex128)


def adjoint_links(U):
    """V_ab = 2 tr(T^a U T^b U^dag), real (N^2-1)^2 per link (dirac.py:358-364)."""
    T = generators(U.shape[1])
    return 2.0 * np.einsum("ijk,mklc,nlo,mjoc->minc", T, U, T, np.conj(U), optimize=True)


def to_adjoint_vector(X):
    """x^a = 2 tr(T^a X) (tests.py:85, dirac.py:531-532)."""
    T = generators(X.shape[1])
    return 2.0 * np.einsum("ijk,mkjc->mic", T, X)


def from_adjoint_vector(x):
    """X = sum_a x^a T^a (hmc.py:204-210)."""
    T = generators(int(round(np.sqrt(x.shape[1] + 1))))
    return np.einsum("mic,ijk->mjkc", x, T)


def gamma5(v, hermitian=True):
    """apply_hermiticity: negate spin rows Ns/2..Ns-1 (dirac.py:649-662)."""
    out = np.array(v, copy=True)
    if hermitian:
        ns = v.shape[0]
        out[ns // 2:] *= -1
    return out


def _down_links(links, tables):
    """links at x-mu: translate(U_mu, mu, -1) (dirac.py:372-378)."""
    return np.stack([olat.translate(links[mu], mu, -1, tables) for mu in range(links.shape[0])])


def apply_simple_fundamental(links, psi, kappa, hermitian, tables):
    """Full 4x4 projectors (dirac.py:503-523)."""
    D = links.shape[0]
    g, _ = gammas(D)
    ns = g.shape[1]
    eye = np.eye(ns)
    down = _down_links(links, tables)
    out = np.array(psi, dtype=np.complex128, copy=True)
    for mu in range(D):
        pp = np.einsum("mn,njc->mjc", eye - g[mu], psi)
        sup = olat.translate(pp, mu, +1, tables)
        hup = np.einsum("ijc,mjc->mic", links[mu], sup)
        pm = np.einsum("mn,njc->mjc", eye + g[mu], psi)
        sdn = olat.translate(pm, mu, -1, tables)
        hdn = np.einsum("jic,mjc->mic", np.conj(down[mu]), sdn)
        out = out + kappa * (hup + hdn)
    return gamma5(out, hermitian)


def apply_projected_fundamental(links, psi, kappa, hermitian, tables):
    """Half-spinor projection + LU-derived reconstruction (dirac.py:458-501)."""
    D = links.shape[0]
    g, _ = gammas(D)
    ns = g.shape[1]
    half = ns // 2
    eye = np.eye(ns)
    down = _down_links(links, tables)
    out = np.array(psi, dtype=np.complex128, copy=True)
    for mu in range(D):
        plus = eye - g[mu]
        minus = eye + g[mu]
        hp = np.einsum("mn,njc->mjc", plus[:half], psi)
        hup = np.einsum("ijc,mjc->mic", links[mu], olat.translate(hp, mu, +1, tables))
        hm = np.einsum("mn,njc->mjc", minus[:half], psi)
        hdn = np.einsum("jic,mjc->mic", np.conj(down[mu]), olat.translate(hm, mu, -1, tables))
        plp, _ = lu(plus, permute_l=True)
        plm, _ = lu(minus, permute_l=True)
        hopping = np.zeros_like(psi)
        hopping[:half] = kappa * (hup + hdn)
        for n1 in range(half, ns):
            for n2 in range(half):
                hopping[n1] += kappa * (plp[n1, n2] * hup[n2] + plm[n1, n2] * hdn[n2])
        out = out + hopping
    return gamma5(out, hermitian)


def apply_simple_adjoint(links, links_orig, X, kappa, hermitian, tables):
    """Adjoint matrix form U X U^dag, full projectors (dirac.py:380-402)."""
    D = links.shape[0]
    g, _ = gammas(D)
    ns = g.shape[1]
    eye = np.eye(ns)
    down = _down_links(links, tables)
    down_o = _down_links(links_orig, tables)
    out = np.array(X, dtype=np.complex128, copy=True)
    for mu in range(D):
        pp = np.einsum("mn,njkc->mjkc", eye - g[mu], X)
        sup = olat.translate(pp, mu, +1, tables)
        hup = np.einsum("ijc,mjkc,lkc->milc", links[mu], sup, np.conj(links_orig[mu]), optimize=True)
        pm = np.einsum("mn,njkc->mjkc", eye + g[mu], X)
        sdn = olat.translate(pm, mu, -1, tables)
        hdn = np.einsum("jic,mjkc,klc->milc", np.conj(down[mu]), sdn, down_o[mu], optimize=True)
        out = out + kappa * (hup + hdn)
    return gamma5(out, hermitian)


def back_projectors(ndim, kappa, hermitian):
    """kappa-folded LU back-projectors (Ns, Ns/2) per mu (dirac.py:302-344)."""
    g, g5 = gammas(ndim)
    ns = g.shape[1]
    half = ns // 2
    eye = np.eye(ns)
    plus_p, minus_p, back_p, back_m = [], [], [], []
    for mu in range(ndim):
        plus = eye - g[mu]
        minus = eye + g[mu]
        if hermitian:
            plus = g5 @ plus
            minus = g5 @ minus
        plus_p.append(plus)
        minus_p.append(minus)
        plp, _ = lu(plus, permute_l=True)
        plm, _ = lu(minus, permute_l=True)
        bp = np.zeros((ns, half), dtype=np.complex128)
        bm = np.zeros((ns, half), dtype=np.complex128)
        for n in range(half):
            bp[n, n] = kappa
            bm[n, n] = kappa
        bp[half:] = kappa * plp[half:, :half]
        bm[half:] = kappa * plm[half:, :half]
        back_p.append(bp)
        back_m.append(bm)
    return np.array(plus_p), np.array(minus_p), np.array(back_p), np.array(back_m)


def apply_projected_adjoint(links, links_orig, X, kappa, hermitian, tables):
    """Adjoint matrix form with half spinors and back-projectors (dirac.py:404-456)."""
    D = links.shape[0]
    ns = gammas(D)[0].shape[1]
    half = ns // 2
    plus_p, minus_p, back_p, back_m = back_projectors(D, kappa, hermitian)
    down = _down_links(links, tables)
    down_o = _down_links(links_orig, tables)
    out = gamma5(X, hermitian)
    for mu in range(D):
        hp = np.einsum("mn,njkc->mjkc", plus_p[mu][:half], X)
        sup = olat.translate(hp, mu, +1, tables)
        hup = np.einsum("ijc,mjkc,lkc->milc", links[mu], sup, np.conj(links_orig[mu]), optimize=True)
        hm = np.einsum("mn,njkc->mjkc", minus_p[mu][:half], X)
        sdn = olat.translate(hm, mu, -1, tables)
        hdn = np.einsum("jic,mjkc,klc->milc", np.conj(down[mu]), sdn, down_o[mu], optimize=True)
        out = out + np.einsum("mn,nijc->mijc", back_p[mu], hup)
        out = out + np.einsum("mn,nijc->mijc", back_m[mu], hdn)
    return out


class WilsonOracle:
    """Bundles links, kappa, BC and representation; ``apply`` = D or Q.

    ``links_orig`` is the gauge field before the boundary provider;
    ``bc_factor`` is the antiperiodic factor on the last direction (None or 1
    for periodic).  For the adjoint representation the matrix form is used.
    """

    def __init__(self, links_orig, geometry, kappa, hermitian=False,
                 representation="fundamental", bc_factor=None, mode="projected"):
        self.geometry = tuple(int(L) for L in geometry)
        self.tables = olat.lookup_tables(self.geometry)
        self.links_orig = np.asarray(links_orig, dtype=np.complex128)
        self.bc_factor = bc_factor
        if bc_factor is None or bc_factor == 1:
            self.links = self.links_orig
        else:
            self.links = olat.apply_antiperiodic(self.links_orig, self.geometry, bc_factor)
        self.kappa = kappa
        self.hermitian = hermitian
        self.representation = representation
        self.mode = mode

    def apply(self, v):
        if self.representation == "fundamental":
            f = apply_projected_fundamental if self.mode == "projected" else apply_simple_fundamental
            return f(self.links, v, self.kappa, self.hermitian, self.tables)
        f = apply_projected_adjoint if self.mode == "projected" else apply_simple_adjoint
        return f(self.links, self.links_orig, v, self.kappa, self.hermitian, self.tables)

    def apply_square(self, v, shift=0.0):
        """SquareOperator then ShiftedOperator (dirac.py:15-37)."""
        return self.apply(self.apply(v)) + shift * v


# ---------------------------------------------------------------------------
# Brute force: the operator assembled element by element from its definition,
# with loops over sites, directions and spin/colour indices.  Shares no code
# with the einsum/gather implementations above.
# ---------------------------------------------------------------------------

def dense_matrix_fundamental(links, geometry, kappa, hermitian):
    geometry = tuple(geometry)
    D = len(geometry)
    g, g5 = gammas(D)
    ns = g.shape[1]
    nc = links.shape[1]
    V = int(np.prod(geometry))
    n = ns * nc * V
    M = np.zeros((n, n), dtype=np.complex128)
    eye = np.eye(ns)

    def idx(s, a, site):
        return (s * nc + a) * V + site

    for site in range(V):
        x = np.unravel_index(site, geometry)
        for s in range(ns):
            for a in range(nc):
                M[idx(s, a, site), idx(s, a, site)] += 1.0
        for mu in range(D):
            xp = list(x)
            xp[mu] = (xp[mu] + 1) % geometry[mu]
            xm = list(x)
            xm[mu] = (xm[mu] - 1) % geometry[mu]
            sp = int(np.ravel_multi_index(xp, geometry))
            sm = int(np.ravel_multi_index(xm, geometry))
            Pp = eye - g[mu]
            Pm = eye + g[mu]
            Uf = links[mu, :, :, site]
            Ub = np.conj(links[mu, :, :, sm]).T
            for s in range(ns):
                for t in range(ns):
                    for a in range(nc):
                        for b in range(nc):
                            M[idx(s, a, site), idx(t, b, sp)] += kappa * Pp[s, t] * Uf[a, b]
                            M[idx(s, a, site), idx(t, b, sm)] += kappa * Pm[s, t] * Ub[a, b]
    if hermitian:
        G = np.kron(g5, np.eye(nc * V))
        M = G @ M
    return M


def free_field_eigenvalue(p, kappa, ndim):
    """Closed form on a plane wave e^{ipx} u with U = 1:
    D = 1 + 2 kappa sum cos p_mu - 2 i kappa sum gamma_mu sin p_mu (a matrix in spin)."""
    g, _ = gammas(ndim)
    ns = g.shape[1]
    M = np.eye(ns, dtype=np.complex128) * (1 + 2 * kappa * np.sum(np.cos(p)))
    for mu in range(ndim):
        M = M - 2j * kappa * np.sin(p[mu]) * g[mu]
    return M
