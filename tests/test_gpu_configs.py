"""Benchmark configurations solved at their full sizes.

The largest single-GPU workload: 4D 32^3 x 64 SU(3),
kappa = 0.1340, near-identity links (eps = 0.2), one Gaussian source, 24
shifts log-spaced in [1e-5, 20] (a rational (D^dag D)^{-1/2}-style shift
list), multishift CG with 16-bit Krylov storage (FP64, then complex64, then
int16 + per-site norms) and FP64 shift solutions, defect correction.

The solver's own FP64 residual check is not trusted here: every shift's true
relative residual ||b - (Q^2 + s) x|| / ||b|| is recomputed from the returned
complex128 solution with the complex128 Wilson operator (whose parity with
the oracle is tested in test_gpu_wilson.py), and must be <= 1e-10.  The
oracle itself is too slow at this size (2M sites); the small-lattice tests of
the same solver against the oracle are in test_gpu_solvers.py.

Also at full size: the 16^3 x 32 12-shift multishift solve (kappa 0.13,
eps 0.3, shifts in [1e-4, 10]) in mixed precision, and a batch of sources
through the batched 16-bit multishift solver at 24^3 x 48 (kappa 0.1335,
eps 0.25, point and Z2 sources, 12 shifts)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def true_residuals(op, b, xs, shifts):
    """||b - (Q^2 + s) x|| / ||b|| per shift with the complex128 operator."""
    bn = torch.linalg.vector_norm(b).item()
    out = []
    for sh, x in zip(shifts, xs):
        assert x.dtype == torch.complex128
        r = b - op.apply(op.apply(x)) - sh * x
        out.append(torch.linalg.vector_norm(r).item() / bn)
    return out


def test_c5_multishift_half_true_residual():
    from tfleonardym_amd import dirac, dirac_solver, fields, lattice
    geometry = (32, 32, 32, 64)
    print("case c5 32^3x64 SU(3) 24 shifts half", flush=True)
    U = fields.gauge_field(geometry, 3, seed=20240, kind="near_identity", eps=0.2)
    b = torch.from_numpy(fields.gaussian_spinor(geometry, 4, 3, seed=20241)).cuda()
    shifts = fields.log_shifts(1e-5, 20.0, 24)
    op = dirac.DiracWilsonOperator(lattice.Configuration(U, geometry, 3, 4), 0.1340, True,
                                   representation="fundamental")
    del U
    s = dirac_solver.multishift_solver(20000, 0.0, algorithm="cg", tolerance=1e-10, precision_mode="half",
                                       verbose=False)
    xs = s.solve(dirac.SquareOperator(op), b, shifts)
    st = s.last_stats
    print(f"iterations {st['iterations']} phase {st['precision_phase']} refine {st['refine_cycles']}", flush=True)
    assert s.last_status == "converged", st
    assert st["precision_phase"] == 2
    bn = torch.linalg.vector_norm(b).item()
    worst = 0.0
    for sh, x, reported in zip(shifts, xs, st["true_rel_resid"]):
        assert x.dtype == torch.complex128
        ax = op.apply(op.apply(x))
        r = b - ax - sh * x
        rel = torch.linalg.vector_norm(r).item() / bn
        worst = max(worst, rel)
        assert rel <= 1e-10, (sh, rel)
        # the solver's own FP64 check agrees with this one
        assert abs(rel - reported) <= 1e-3 * max(rel, reported) + 1e-14, (sh, rel, reported)
    print(f"worst true relative residual {worst:.3e}", flush=True)


def test_c2_multishift_mixed_true_residual():
    from tfleonardym_amd import dirac, dirac_solver, fields, lattice
    geometry = (16, 16, 16, 32)
    print("case c2 16^3x32 SU(3) 12 shifts mixed", flush=True)
    U = fields.gauge_field(geometry, 3, seed=20240, kind="near_identity", eps=0.3)
    b = torch.from_numpy(fields.gaussian_spinor(geometry, 4, 3, seed=20241)).cuda()
    shifts = fields.log_shifts(1e-4, 10.0, 12)
    op = dirac.DiracWilsonOperator(lattice.Configuration(U, geometry, 3, 4), 0.13, True,
                                   representation="fundamental")
    s = dirac_solver.multishift_solver(20000, 0.0, algorithm="cg", tolerance=1e-10, precision_mode="mixed",
                                       verbose=False)
    xs = s.solve(dirac.SquareOperator(op), b, shifts)
    assert s.last_status == "converged", s.last_stats
    rel = true_residuals(op, b, xs, shifts)
    assert max(rel) <= 1e-10, rel


def test_c3_batched_multishift_half_true_residual():
    from tfleonardym_amd import dirac, dirac_solver, fields, lattice
    geometry = (24, 24, 24, 48)
    print("case c3 24^3x48 SU(3) batched multishift half", flush=True)
    U = fields.gauge_field(geometry, 3, seed=20250, kind="near_identity", eps=0.25)
    srcs = [fields.point_source(geometry, 4, 3, site=0, color=0, spin=0),
            fields.point_source(geometry, 4, 3, site=0, color=2, spin=3),
            fields.z2_noise(geometry, 4, 3, seed=20260), fields.z2_noise(geometry, 4, 3, seed=20261)]
    srcs = [torch.from_numpy(x).cuda() for x in srcs]
    shifts = fields.log_shifts(1e-4, 10.0, 12)
    op = dirac.DiracWilsonOperator(lattice.Configuration(U, geometry, 3, 4), 0.1335, True,
                                   representation="fundamental")
    s = dirac_solver.multishift_solver(20000, 0.0, algorithm="cg", tolerance=1e-10, precision_mode="half",
                                       verbose=False)
    sols = s.solve_batch(dirac.SquareOperator(op), srcs, shifts)
    assert s.last_status == "converged", s.last_stats
    assert all(st["precision_phase"] == 2 for st in s.last_stats)
    for b, row, st in zip(srcs, sols, s.last_stats):
        rel = true_residuals(op, b, row, shifts)
        assert max(rel) <= 1e-10, rel
        for r, reported in zip(rel, st["true_rel_resid"]):
            assert abs(r - reported) <= 1e-3 * max(r, reported) + 1e-14
