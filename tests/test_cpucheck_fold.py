"""CPU check of the folded deferred multishift update (fold_shift_block in
csrc/lym_fin.h, compiled for the host) against the step-by-step CG-M shift
recurrence it replaces (Jegerlehner, hep-lat/9612014; the shift updates of
oracle/solvers.py's multishift CG): for every step j < mk,
    x += alpha_j p ;  p = zeta_j r_j + beta_j p ,
with r_j the base residual of step j.  Covers a full block, a shift that
stopped inside the block (mk < nb), one step, and a block of the largest
length."""
import ctypes
import os

import numpy as np
import pytest

LIB = os.path.join(os.path.dirname(__file__), "..", "tfleonardym_amd", "liblym_cpucheck.so")
MAX_BLOCK = 128


@pytest.fixture(scope="module")
def lib():
    from tfleonardym_amd import build
    build.build_cpu_check()
    lib = ctypes.CDLL(LIB)
    lib.lymcheck_fold_shift_block.restype = ctypes.c_int
    lib.lymcheck_fold_shift_block.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p]
    return lib


def stepwise(a, z, b, x, p, r, mk):
    x = x.copy()
    p = p.copy()
    for j in range(mk):
        x = x + a[j] * p
        p = z[j] * r[j] + b[j] * p
    return x, p


def folded(lib, a, z, b, x, p, r, mk):
    nb = r.shape[0]
    n = x.size
    xo = np.ascontiguousarray(x.astype(np.complex128))
    po = np.ascontiguousarray(p.astype(np.complex128))
    rr = np.ascontiguousarray(r.astype(np.complex128))
    a = np.ascontiguousarray(a, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    rc = lib.lymcheck_fold_shift_block(mk, nb, a.ctypes.data, z.ctypes.data, b.ctypes.data, n, rr.ctypes.data,
                                       xo.ctypes.data, po.ctypes.data)
    assert rc == 0
    return xo, po


def cgm_like_coefficients(rng, m):
    """Step coefficients of the sizes a converging shift produces: alpha_s
    ~ 1, zeta_s falling, beta_s = beta (zeta_{k+1} / zeta_k)^2 below 1."""
    beta = rng.uniform(0.6, 0.99, m)
    zeta = np.cumprod(rng.uniform(0.8, 1.0, m))
    a = rng.uniform(0.3, 1.5, m) * zeta
    b = beta * np.concatenate([[1.0], (zeta[1:] / zeta[:-1]) ** 2])
    return a, zeta, b


@pytest.mark.parametrize("mk,nb", [(16, 16), (64, 64), (40, 64), (1, 1), (1, 32), (MAX_BLOCK, MAX_BLOCK)])
def test_fold_matches_stepwise(lib, mk, nb):
    rng = np.random.default_rng(1000 + 7 * mk + nb)
    n = 96
    a, z, b = cgm_like_coefficients(rng, nb)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    p = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r = rng.standard_normal((nb, n)) + 1j * rng.standard_normal((nb, n))
    xe, pe = stepwise(a, z, b, x, p, r, mk)
    xf, pf = folded(lib, a, z, b, x, p, r, mk)
    assert np.linalg.norm(xf - xe) <= 1e-13 * np.linalg.norm(xe)
    assert np.linalg.norm(pf - pe) <= 1e-13 * np.linalg.norm(pe)


def test_fold_no_steps_is_identity(lib):
    rng = np.random.default_rng(5)
    n = 8
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    p = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r = rng.standard_normal((4, n)) + 1j * rng.standard_normal((4, n))
    xf, pf = folded(lib, np.ones(4), np.ones(4), np.ones(4), x, p, r, 0)
    assert np.array_equal(xf, x) and np.array_equal(pf, p)


def test_fold_single_step_closed_form(lib):
    # one step: x + a p and z r + b p exactly (no rounding beyond one FMA)
    x = np.array([1.0 + 2.0j]), np.array([0.5 - 1.0j])
    r = np.array([[3.0 + 0.25j]])
    xf, pf = folded(lib, np.array([2.0]), np.array([4.0]), np.array([0.5]), x[0], x[1], r, 1)
    assert xf[0] == 1.0 + 2.0j + 2.0 * (0.5 - 1.0j)
    assert pf[0] == 4.0 * (3.0 + 0.25j) + 0.5 * (0.5 - 1.0j)


def test_fold_rejects_bad_lengths(lib):
    z = np.zeros(2)
    v = np.zeros(2, dtype=np.complex128)
    assert lib.lymcheck_fold_shift_block(3, 2, z.ctypes.data, z.ctypes.data, z.ctypes.data, 1, v.ctypes.data,
                                         v.ctypes.data, v.ctypes.data) != 0
    assert lib.lymcheck_fold_shift_block(1, MAX_BLOCK + 1, z.ctypes.data, z.ctypes.data, z.ctypes.data, 1,
                                         v.ctypes.data, v.ctypes.data, v.ctypes.data) != 0
