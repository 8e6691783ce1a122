"""Random inputs of the fermion path, drop-in for hmc.random_gauge_field,
hmc.hotstart and hmc.generate_random_vector (hmc.py:16-31, 196-210).

Gaussian draws come from torch's device generator (optionally seeded; the
reference draws from TensorFlow's global generator), the SU(N) projection
from the lym_unitary_projection kernel.  The volume is the installed lookup
tables' local volume (lattice.set_translate), as in the reference; hotstart
falls back to the volume of `geometry` when no tables are installed.  The
HMC integrators of that module are outside the fermion path."""
import numpy as np
import torch

from . import _lib
from . import lattice
from . import lie_generators
from . import smearing


def _volume(volume=None):
    if volume is not None:
        return int(volume)
    if lattice.lookup_tables is None:
        raise RuntimeError("call lattice.set_translate(False, geometry) first (or pass volume)")
    return int(lattice.lookup_tables.local_volume)


def _normal(shape, stddev, generator):
    dev = _lib.require_device()
    re = torch.randn(shape, dtype=torch.float64, device=dev, generator=generator) * stddev
    im = torch.randn(shape, dtype=torch.float64, device=dev, generator=generator) * stddev
    return torch.complex(re, im)


def random_gauge_field(nc=3, nd=4, volume=None, generator=None):
    """Gaussian complex matrices projected onto SU(nc) (hmc.py:16-24)."""
    return smearing.unitary_sqrt_projection(_normal((nd, nc, nc, _volume(volume)), 1.0, generator))


def hotstart(geometry, nc=3, nd=4, generator=None):
    """Configuration of random SU(nc) links (hmc.py:27-31)."""
    volume = None if lattice.lookup_tables is not None else int(np.prod(geometry))
    return lattice.Configuration(gauge_field=random_gauge_field(nc, nd, volume, generator), geometry=geometry,
                                 colors=nc, number_of_dimensions=nd)


def generate_random_vector(nc=3, spinor_dimension=4, representation="fundamental", stddev=1, volume=None,
                           generator=None):
    """Gaussian spinor, re and im ~ N(0, stddev^2) (hmc.py:196-210); adjoint
    vectors are the matrices sum_a x^a T^a."""
    V = _volume(volume)
    if representation == "fundamental":
        return _normal((spinor_dimension, nc, V), stddev, generator)
    if representation == "adjoint":
        x = _normal((spinor_dimension, nc * nc - 1, V), stddev, generator)
        T = torch.from_numpy(lie_generators.generators(nc, "fundamental")).to(x.device)
        return torch.einsum("mic,ijk->mjkc", x, T).contiguous()
    raise RuntimeError(f"unknown representation {representation}")
