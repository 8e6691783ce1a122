"""Pins the oracle's solvers (dirac_solver.py:13-110 and textbook CG / CG-M)
against numpy.linalg.solve on dense tiny-lattice operators and the
reference's own invariants (tests.py:96-119)."""
import numpy as np
import pytest

from oracle import solvers as osv
from oracle import wilson as ow
from tfleonardym_amd import fields


@pytest.fixture(scope="module")
def system():
    geometry = (4, 4)
    U = fields.gauge_field(geometry, 2, seed=5, kind="haar")
    op = ow.WilsonOracle(U, geometry, 0.15, hermitian=True, bc_factor=-1.0)
    M = ow.dense_matrix_fundamental(op.links, geometry, 0.15, True)
    A = M @ M
    b = fields.gaussian_spinor(geometry, 2, 2, seed=1)
    return op, A, b


def direct(A, b, shift):
    n = A.shape[0]
    return np.linalg.solve(A + shift * np.eye(n), b.reshape(-1)).reshape(b.shape)


def test_cg_matches_direct(system):
    op, A, b = system
    x, it = osv.cg(op.apply_square, b, shift=0.01, tol=1e-12)
    assert np.linalg.norm(x - direct(A, b, 0.01)) / np.linalg.norm(direct(A, b, 0.01)) < 1e-9
    assert osv.true_residual(op.apply_square, b, x, 0.01) < 1e-11


def test_cgm_matches_direct(system):
    op, A, b = system
    shifts = fields.log_shifts(1e-3, 5.0, 7)
    xs, it = osv.cgm(op.apply_square, b, shifts, tol=1e-12)
    for s, x in zip(shifts, xs):
        want = direct(A, b, s)
        assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-9, s
        assert osv.true_residual(op.apply_square, b, x, s) < 1e-11
    # single-shift CG-M is plain CG iteration for iteration
    x1, it1 = osv.cgm(op.apply_square, b, [shifts[-1]], tol=1e-12)
    x2, it2 = osv.cg(op.apply_square, b, shift=shifts[-1], tol=1e-12)
    assert it1 == it2
    assert np.allclose(x1[0], x2, rtol=0, atol=1e-12 * np.abs(x2).max())


def test_mr_multishift_residuals(system):
    """tests.py:96-106: each shifted solution satisfies its system."""
    op, A, b = system
    shifts = [0.5, 0.1, 0.01]
    xs, step, ok = osv.mr_multishift(op.apply_square, b, shifts, 1000, 1e-11)
    assert ok
    bb = np.vdot(b, b).real
    for s, x in zip(shifts, xs):
        r = op.apply_square(x, s) - b
        # the base shift's iterated residual is below 1e-11 absolute; the
        # other shifts' residuals are f_i-scaled copies, hence smaller
        assert np.vdot(r, r).real < 1e-9 * bb
        want = direct(A, b, s)
        assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-5


def test_bicgstab_and_mr_agree(system):
    """tests.py:108-113: BiCGStab on the last shift agrees with MR."""
    op, A, b = system
    xs, _, _ = osv.mr_multishift(op.apply_square, b, [0.5, 0.1, 0.01], 1000, 1e-11)
    x, step, ok = osv.bicgstab(lambda v: op.apply_square(v, 0.01), b, 1000, 1e-11)
    assert ok
    want = direct(A, b, 0.01)
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-6
    assert np.linalg.norm(x - xs[-1]) / np.linalg.norm(want) < 1e-5


def test_mr_negative_leading_shift_uses_bicgstab(system):
    op, A, b = system
    shifts = [-50.0, 0.3, 0.05]
    xs, step, ok = osv.mr_multishift(op.apply_square, b, shifts, 2000, 1e-12)
    for s, x in zip(shifts, xs):
        want = direct(A, b, s)
        assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-5


def test_inverse_square_is_hermitian(system):
    """tests.py:115-119: <v2, A^-1 v1> = <A^-1 v2, v1>."""
    op, A, b = system
    v2 = fields.gaussian_spinor((4, 4), 2, 2, seed=8)
    x1, _ = osv.cg(op.apply_square, b, 0.0, tol=1e-13)
    x2, _ = osv.cg(op.apply_square, v2, 0.0, tol=1e-13)
    a = np.vdot(v2, x1)
    c = np.vdot(x2, b)
    assert abs(a - c) < 1e-9 * abs(a)
