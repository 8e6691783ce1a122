"""The torch-CPU restatement used as the benchmark's CPU baseline agrees with
the NumPy oracle: the Wilson operator element for element, and the multishift
CG iterations (base residual after a fixed number of iterations against the
true residual of the oracle's CG-M solution after as many)."""
import numpy as np
import pytest
import torch

from oracle import solvers as osv
from oracle import torch_cpu as otc
from oracle import wilson as ow
from tfleonardym_amd import fields


@pytest.mark.parametrize("geometry,nc,herm", [((4, 4, 4, 4), 3, True), ((2, 4, 4, 6), 3, False),
                                              ((6, 4), 2, True), ((4, 4, 2, 4), 2, False)])
def test_torch_wilson_matches_oracle(geometry, nc, herm):
    U = fields.gauge_field(geometry, nc, seed=5, kind="haar")
    ns = 4 if len(geometry) == 4 else 2
    psi = fields.gaussian_spinor(geometry, ns, nc, seed=6)
    want = ow.WilsonOracle(U, geometry, 0.13, herm, "fundamental").apply(psi)
    got = otc.TorchWilson(U, geometry, 0.13, herm).apply(torch.from_numpy(psi)).numpy()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-13


def test_torch_cgm_iterations_match_oracle():
    geometry = (4, 4, 4, 4)
    U = fields.gauge_field(geometry, 3, seed=7, kind="near_identity", eps=0.3)
    b = fields.gaussian_spinor(geometry, 4, 3, seed=8)
    shifts = [1.0, 0.1, 0.01]
    op = otc.TorchWilson(U, geometry, 0.13, True)
    rn = otc.cgm_iterations(op, torch.from_numpy(b), shifts, 6)
    oop = ow.WilsonOracle(U, geometry, 0.13, True, "fundamental")
    xs, _ = osv.cgm(oop.apply_square, b, shifts, tol=1e-30, max_iter=6)
    r = b - oop.apply_square(xs[2], 0.01)
    assert abs(rn - np.linalg.norm(r)) <= 1e-10 * np.linalg.norm(b)
    assert rn < 0.5 * np.linalg.norm(b)
