"""Lattice geometry: the Configuration record, neighbour tables and translate.

Mirrors lattice.py:7 (``Configuration``), lattice.py:86-128
(``mth_lookup_tables``, ``mth_translate``, ``mth_global_sum``) and
translate.py:9-36 (``set_translate``) of the reference.  Sites are numbered
with the last coordinate fastest.  The Wilson kernels compute neighbours from
coordinates and never read these tables; they serve callers off the operator
path (e.g. ``translate`` of a field by one site).
"""
import collections
import functools

import numpy as np
import torch

from . import _lib

Configuration = collections.namedtuple(
    "Configuration", ["gauge_field", "geometry", "colors", "number_of_dimensions"])

MthLookupTables = collections.namedtuple(
    "MthLookupTables", ["sup", "sdn", "global_coordinates", "global_volume", "local_volume"])


@functools.lru_cache(maxsize=100)
def _tables_np(geometry):
    D = len(geometry)
    grids = np.meshgrid(*[np.arange(L) for L in geometry], indexing="ij")
    coords = np.stack([g.reshape(-1) for g in grids]).astype(np.int64)
    sup = np.empty((D, coords.shape[1]), dtype=np.int64)
    sdn = np.empty((D, coords.shape[1]), dtype=np.int64)
    for mu in range(D):
        for sign, out in ((1, sup), (-1, sdn)):
            c = coords.copy()
            c[mu] = (c[mu] + sign) % geometry[mu]
            s = np.zeros(coords.shape[1], dtype=np.int64)
            for nu in range(D):
                s = s * geometry[nu] + c[nu]
            out[mu] = s
    return sup, sdn, coords


def mth_lookup_tables(lattice_geometry):
    """sup/sdn (D, V) int64 neighbour tables and (D, V) coordinates, on the device
    (lattice.py:86-120)."""
    g = tuple(int(L) for L in lattice_geometry)
    sup, sdn, coords = _tables_np(g)
    dev = _lib.require_device()
    V = int(np.prod(g))
    return MthLookupTables(torch.from_numpy(sup).to(dev), torch.from_numpy(sdn).to(dev),
                           torch.from_numpy(coords).to(dev), V, V)


lookup_tables = None


def mth_translate(tensor, direction, sign, lookup_tables):
    """translate(T, mu, +1)[..., s] = T[..., sup[mu][s]] (lattice.py:124-128)."""
    idx = lookup_tables.sup[direction] if sign > 0 else lookup_tables.sdn[direction]
    return torch.index_select(tensor, tensor.dim() - 1, idx)


def translate(tensor, direction, sign):
    if lookup_tables is None:
        raise RuntimeError("call set_translate(False, geometry, mpi_grid) first")
    return mth_translate(tensor, direction, sign, lookup_tables)


def mth_global_sum(to_sum):
    """Single-process global sum is the identity (lattice.py:375-376)."""
    return to_sum


global_sum = mth_global_sum


def set_translate(use_mpi, geometry, mpi_grid=None):
    """Install the single-process translate / tables / global_sum
    (translate.py:9-36).  Domain decomposition over several GPUs is not
    implemented: one process drives one GPU."""
    global lookup_tables
    if use_mpi:
        raise NotImplementedError("multi-process domain decomposition is not built (one GPU)")
    lookup_tables = mth_lookup_tables(tuple(geometry))
    return lookup_tables
