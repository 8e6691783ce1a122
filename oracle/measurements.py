"""Oracle: fermion measurements (test infrastructure only).

NumPy restatement of fermion_measurements.py:41-115 and Propagator
(dirac.py:665-696): sol = g5 Q (Q^2)^-1 src with the solve done by CG to a
relative residual of 1e-13, the pion correlator's per-slice density sums
by explicit loops over site coordinates, and the chiral condensate's
mean <eta, sol> / V.

Pinning (tests/test_oracle_measurements.py): Q g5 sol = src (the
propagator inverts Q), slice sums against a dense per-coordinate mask and
their total against |sol|^2, the free-field correlator's reflection
symmetry corr[z] = corr[L - z], and the condensate against tr(g5 Q^-1)
on a dense matrix for a complete set of sources.
"""
import numpy as np

from . import lattice as olat
from . import solvers as osol
from . import wilson as ow


def propagator(wo, src, tol=1e-13):
    """g5 Op (Op^2)^-1 src (dirac.py:671-685)."""
    A = lambda v: wo.apply(wo.apply(v))
    x, _ = osol.cg(A, np.asarray(src, dtype=np.complex128), tol=tol)
    return ow.gamma5(wo.apply(x), True)


def site_density(v, representation):
    """sum_{s,c} |v|^2 or sum_{s,c,d} conj(v_scd) v_sdc per site
    (fermion_measurements.py:75-78)."""
    v = np.asarray(v)
    if representation == "fundamental":
        return np.einsum("icm,icm->m", np.conj(v), v)
    return np.einsum("icdm,idcm->m", np.conj(v), v)


def slice_sums(solutions, geometry, direction, representation):
    """corr[z] = sum over solutions and sites with x_direction = z
    (fermion_measurements.py:73-84)."""
    coords = olat.coordinates(tuple(geometry))
    L = geometry[direction]
    corr = np.zeros(L, dtype=np.complex128)
    for v in solutions:
        d = site_density(v, representation)
        for z in range(L):
            corr[z] += d[coords[direction] == z].sum()
    return corr


def point_sources(ns, nc, V, representation):
    """Ns x (Nc or Nc^2-1) unit sources at site 0, spin-major
    (fermion_measurements.py:61-70)."""
    out = []
    ncol = nc * nc - 1 if representation == "adjoint" else nc
    T = ow.generators(nc) if representation == "adjoint" else None
    for spin in range(ns):
        for c in range(ncol):
            if representation == "fundamental":
                v = np.zeros((ns, nc, V), dtype=np.complex128)
                v[spin, c, 0] = 1.0
            else:
                v = np.zeros((ns, nc, nc, V), dtype=np.complex128)
                v[spin, :, :, 0] = T[c]
            out.append(v)
    return out


def pion_correlator(wo, direction):
    D = wo.links_orig.shape[0]
    ns = 2 if D == 2 else 4
    nc = wo.links_orig.shape[1]
    V = wo.links_orig.shape[-1]
    sols = [propagator(wo, s) for s in point_sources(ns, nc, V, wo.representation)]
    return slice_sums(sols, wo.geometry, direction, wo.representation)


def chiral_condensate(wo, sources):
    V = wo.links_orig.shape[-1]
    return np.mean([np.vdot(s, propagator(wo, s)) / V for s in sources])
