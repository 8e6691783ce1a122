"""Oracle: lattice geometry (test infrastructure only).

Restates lattice.py:86-128 (``mth_lookup_tables``, ``mth_translate``) and
boundary_conditions.py:10-46 in NumPy.  Site index: last coordinate runs
fastest, s = ((x0*L1 + x1)*L2 + x2)*L3 + x3 (lattice.py:92-114).
"""
import numpy as np


def coordinates(geometry):
    """(D, V) int64 coordinates of every site (lattice.py:92-98)."""
    geometry = tuple(int(L) for L in geometry)
    grids = np.meshgrid(*[np.arange(L) for L in geometry], indexing="ij")
    return np.stack([g.reshape(-1) for g in grids]).astype(np.int64)


def site_index(coords, geometry):
    """Horner site index with the last coordinate fastest (lattice.py:103-114)."""
    s = np.zeros(coords.shape[1:], dtype=np.int64)
    for mu, L in enumerate(geometry):
        s = s * L + (coords[mu] % L)
    return s


def lookup_tables(geometry):
    """sup/sdn (D, V) neighbour tables: sup[mu][s] = site of x+mu (lattice.py:86-120)."""
    c = coordinates(geometry)
    D = len(geometry)
    sup = np.empty((D, c.shape[1]), dtype=np.int64)
    sdn = np.empty((D, c.shape[1]), dtype=np.int64)
    for mu in range(D):
        cu = c.copy()
        cu[mu] = (cu[mu] + 1) % geometry[mu]
        sup[mu] = site_index(cu, geometry)
        cd = c.copy()
        cd[mu] = (cd[mu] - 1) % geometry[mu]
        sdn[mu] = site_index(cd, geometry)
    return sup, sdn


def translate(tensor, mu, sign, tables):
    """translate(T, mu, +1)[..., s] = T[..., sup[mu][s]] (lattice.py:124-128)."""
    sup, sdn = tables
    idx = sup[mu] if sign > 0 else sdn[mu]
    return np.take(tensor, idx, axis=-1)


def antiperiodic_field(geometry, factor=-1.0):
    """Per-site factor: ``factor`` where x_{D-1} = L_{D-1}-1, else 1
    (boundary_conditions.py:10-20)."""
    c = coordinates(geometry)
    f = np.ones(c.shape[1], dtype=np.complex128)
    f[c[-1] == geometry[-1] - 1] = factor
    return f


def apply_antiperiodic(gauge_field, geometry, factor=-1.0):
    """Multiply U_{D-1} by the antiperiodic field (boundary_conditions.py:23-46)."""
    out = np.array(gauge_field, dtype=np.complex128, copy=True)
    out[-1] = out[-1] * antiperiodic_field(geometry, factor)[None, None, :]
    return out
