"""Fermion measurements, drop-in for the reference's fermion_measurements
(fermion_measurements.py:11-115): Z2-noise and point sources, the pion
correlator and the chiral condensate.

The propagators are the reference's ``Propagator`` (dirac.py:665-696):
sol = g5 Op (Op^2)^-1 src.  By default every source is solved as the
reference does, with BiCGStab to its absolute precision (5000 steps; 1e-13
for the correlator, 1e-10 for the condensate).  ``solver="batched"`` solves
all sources of a measurement together in the batched multishift CG (one
Wilson launch per iteration covers every source; complex64 Krylov vectors
with FP64 reliable updates) to the true relative residual ``tolerance``.
The per-slice sums of the correlator run in the lym_slice_sums kernels; the
condensate's inner products in the device dot kernel.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import algebra_utils
from . import dirac
from . import dirac_solver
from . import lie_generators
from .measurements import Measurement
from .tensors import to_device


def _rng(generator):
    if generator is None:
        return np.random.default_rng()
    if isinstance(generator, np.random.Generator):
        return generator
    return np.random.default_rng(generator)


def generate_random_z2_noise(volume, nc=3, spinor_dimension=4, representation="fundamental", generator=None):
    """+-1 per component (fermion_measurements.py:11-23); adjoint sources are
    the matrices sum_a x^a T^a.  `generator`: a numpy Generator or seed (the
    reference draws from TensorFlow's global generator)."""
    rng = _rng(generator)
    if representation == "fundamental":
        v = (rng.integers(0, 2, (spinor_dimension, nc, volume)) * 2.0 - 1.0).astype(np.complex128)
        return to_device(v)
    if representation == "adjoint":
        x = (rng.integers(0, 2, (spinor_dimension, nc * nc - 1, volume)) * 2.0 - 1.0).astype(np.complex128)
        return to_device(np.einsum("mic,ijk->mjkc", x, lie_generators.generators(nc, "fundamental")))
    raise RuntimeError(f"unknown representation {representation}")


def generate_point_source(site, color, spin, volume, nc=3, spinor_dimension=4, representation="fundamental"):
    """Unit source at (spin, color, site) (fermion_measurements.py:26-38); for
    the adjoint, colour indexes the generator: the source is T^color at the site."""
    if representation == "fundamental":
        v = np.zeros((spinor_dimension, nc, volume), dtype=np.complex128)
        v[spin, color, site] = 1.0
        return to_device(v)
    if representation == "adjoint":
        v = np.zeros((spinor_dimension, nc, nc, volume), dtype=np.complex128)
        v[spin, :, :, site] = lie_generators.generators(nc, "fundamental")[color]
        return to_device(v)
    raise RuntimeError(f"unknown representation {representation}")


def propagate(dirac_operator, sources, solver=None, precision=1e-13, tolerance=1e-10, max_steps=5000):
    """g5 Op (Op^2)^-1 src for every source (Propagator.apply, dirac.py:671-685).
    solver: None -> BiCGStab(max_steps, precision) per source as the reference
    measurements build it; "batched" -> all sources in one batched CG to the
    true relative residual `tolerance` (hermitian Wilson operators)."""
    if solver is None:
        prop = dirac.Propagator(dirac_operator, dirac_solver.biconjugate_gradient(max_steps, precision,
                                                                                  verbose=False))
        return [prop.apply(s) for s in sources]
    if solver != "batched":
        raise RuntimeError(f"unknown propagator solver {solver}")
    if not (isinstance(dirac_operator, dirac.DiracWilsonOperator) and dirac_operator.hermitian):
        raise RuntimeError("the batched propagator needs a hermitian DiracWilsonOperator")
    ms = dirac_solver.multishift_solver(max(max_steps, 20000), 0.0, algorithm="cg", tolerance=tolerance,
                                        precision_mode="mixed", verbose=False)
    xs = ms.solve_batch(dirac.SquareOperator(dirac_operator), sources, [0.0])
    prop = dirac.Propagator(dirac_operator, None)
    return [prop.apply_gammad(dirac_operator.apply(x[0])) for x in xs]


def slice_sums(dirac_operator, solutions, direction):
    """Per-slice sums of the pion density over `solutions` (lym_slice_sums):
    a complex128 array of length L_direction (the values are real)."""
    op = dirac_operator
    L = op.geometry[direction]
    vs = [to_device(v) for v in solutions]
    shape = tuple(op.spinor_shape)
    for v in vs:
        if tuple(v.shape) != shape:
            raise RuntimeError(f"solution shape {tuple(v.shape)} does not match the operator's {shape}")
    dev = _lib.require_device()
    scratch = torch.empty(op.volume + L, dtype=torch.float64, device=dev)
    out = (ctypes.c_double * L)()
    ptrs = (ctypes.c_void_p * max(len(vs), 1))(*[v.data_ptr() for v in vs])
    _lib.check(_lib.lib().lym_slice_sums(ctypes.byref(op.desc), int(direction), len(vs), ptrs, _lib.ptr(scratch),
                                         scratch.numel() * 8, out, _lib.stream_handle(dev)), "slice_sums")
    return np.array([complex(out[z], 0.0) for z in range(L)])


def pion_correlator(dirac_operator, correlator_direction, lookup_tables=None, global_summator=None, solver=None,
                    tolerance=1e-10):
    """fermion_measurements.py:41-92: Ns * Nc point sources at site 0 (Nc^2-1
    generators for the adjoint), their propagators, and per slice along
    `correlator_direction` the summed density.  `lookup_tables` is accepted
    for the reference's signature (the slices come from the lattice geometry).
    `tolerance`: true relative residual of the batched solver."""
    op = dirac_operator
    ncol = op.number_of_colors ** 2 - 1 if op.representation == "adjoint" else op.number_of_colors
    sources = [generate_point_source(0, c, spin, op.volume, op.number_of_colors, op.spinor_dimension,
                                     op.representation)
               for spin in range(op.spinor_dimension) for c in range(ncol)]
    solutions = propagate(op, sources, solver, precision=1e-13, tolerance=tolerance)
    correlator = slice_sums(op, solutions, correlator_direction)
    if global_summator is not None:
        for i in range(len(correlator)):
            correlator[i] = global_summator(correlator[i])
    return Measurement(name="pion_operator", value=np.array(correlator), geometry=op.cfg.geometry,
                       colors=op.cfg.colors)


def chiral_condensate(dirac_operator, num_estimators=20, solver=None, generator=None, tolerance=1e-10):
    """fermion_measurements.py:95-115: mean over Z2 sources eta of
    <eta, g5 Op (Op^2)^-1 eta> / V."""
    op = dirac_operator
    rng = _rng(generator)
    sources = [generate_random_z2_noise(op.volume, op.number_of_colors, op.spinor_dimension, op.representation,
                                        generator=rng) for _ in range(num_estimators)]
    solutions = propagate(op, sources, solver, precision=1e-10, tolerance=tolerance)
    value = np.mean([algebra_utils.dot(s, x) / op.volume for s, x in zip(sources, solutions)])
    return Measurement(name="chiral_condensate", value=value, geometry=op.cfg.geometry, colors=op.cfg.colors)
