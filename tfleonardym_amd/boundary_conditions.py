"""Antiperiodic boundary condition in the last lattice direction.

Mirrors boundary_conditions.py:10-46: ``set_antiperiodic_field_in_t_direction``
builds the per-site factor (``factor`` on the last time slice, 1 elsewhere);
``apply_antiboundary_conditions_in_t_direction`` multiplies U_{D-1} by it and
accepts a ``Configuration`` or a bare (D, Nc, Nc, V) tensor.

When this provider is handed to ``DiracWilsonOperator``, the operator does
not materialise a signed copy of the links: the Wilson kernels apply the
factor on the fly to the links that cross the last slice.
"""
import numpy as np
import torch

from . import _lib
from .lattice import Configuration

antiperiodic_field_in_t_direction = None
_geometry = None
_factor = None


def set_antiperiodic_field_in_t_direction(lattice_geometry, factor=-1):
    global antiperiodic_field_in_t_direction, _geometry, _factor
    g = tuple(int(L) for L in lattice_geometry)
    V = int(np.prod(g))
    last = np.arange(V) % g[-1]
    f = np.ones(V, dtype=np.complex128)
    f[last == g[-1] - 1] = factor
    antiperiodic_field_in_t_direction = torch.from_numpy(f).to(_lib.require_device())
    _geometry = g
    _factor = complex(factor)


def current_factor():
    """(geometry, factor) installed by set_antiperiodic_field_in_t_direction."""
    if _geometry is None:
        raise RuntimeError("call set_antiperiodic_field_in_t_direction(geometry) first")
    return _geometry, _factor


def _twist(field):
    f = antiperiodic_field_in_t_direction
    if f is None:
        raise RuntimeError("call set_antiperiodic_field_in_t_direction(geometry) first")
    field = field if isinstance(field, torch.Tensor) else torch.as_tensor(np.asarray(field))
    field = field.to(device=f.device, dtype=torch.complex128)
    out = field.clone()
    out[-1] = out[-1] * f
    return out


def apply_antiboundary_conditions_in_t_direction(cfg):
    if isinstance(cfg, Configuration):
        return Configuration(gauge_field=_twist(cfg.gauge_field), geometry=cfg.geometry,
                             colors=cfg.colors, number_of_dimensions=cfg.number_of_dimensions)
    return _twist(cfg)
