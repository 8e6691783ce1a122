"""Pins the oracle's SU(N) projection (smearing.py:18-31): the polar factor
is unitary with P^dag X Hermitian positive definite (which makes it unique),
and the result is special unitary and equal to P / det(P)^(1/N)."""
import numpy as np
import pytest

from oracle import hmc as oh


@pytest.mark.parametrize("nc,nd", [(2, 2), (3, 4)])
def test_polar_factor_and_projection(nc, nd):
    rng = np.random.default_rng(3)
    V = 50
    X = rng.standard_normal((nd, nc, nc, V)) + 1j * rng.standard_normal((nd, nc, nc, V))
    P = oh.polar_factor(X)
    W = oh.unitary_sqrt_projection(X)
    eye = np.eye(nc)
    for mu in range(nd):
        for s in range(V):
            p, x, w = P[mu, :, :, s], X[mu, :, :, s], W[mu, :, :, s]
            assert np.allclose(p.conj().T @ p, eye, atol=1e-12)
            h = p.conj().T @ x
            assert np.allclose(h, h.conj().T, atol=1e-11)
            assert np.linalg.eigvalsh((h + h.conj().T) / 2).min() > 0
            assert np.allclose(w.conj().T @ w, eye, atol=1e-12)
            assert abs(np.linalg.det(w) - 1) < 1e-12
            d = np.linalg.det(p)
            r = abs(d) ** (1.0 / nc) * np.exp(1j * np.angle(d) / nc)
            assert np.allclose(w * r, p, atol=1e-12)
