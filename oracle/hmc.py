"""Oracle: SU(N) projection of random link matrices (test infrastructure only).

NumPy restatement of unitary_sqrt_projection (smearing.py:18-31), which
hotstart uses (hmc.py:16-31): U = X (X^dag X)^-1/2 per link, with the
inverse square root of the Hermitian positive matrix from its eigen-
decomposition, divided by det(U)^(1/N) on the principal branch.

Pinning (tests/test_oracle_hmc.py): the polar factor is the unique unitary P
with P^dag X Hermitian positive definite; the result is P / det(P)^(1/N),
special unitary.
"""
import numpy as np


def polar_factor(X):
    """(D, N, N, V) -> the unitary polar factor of every link."""
    D, N, _, V = X.shape
    M = np.asarray(X).transpose(0, 3, 1, 2).reshape(-1, N, N)
    H = np.einsum("lmk,lmn->lkn", np.conj(M), M)
    lam, W = np.linalg.eigh(H)
    Hinv_sqrt = np.einsum("lij,lj,lkj->lik", W, lam ** -0.5, np.conj(W))
    P = np.einsum("lmk,lkn->lmn", M, Hinv_sqrt)
    return P.reshape(D, V, N, N).transpose(0, 2, 3, 1)


def unitary_sqrt_projection(X):
    """smearing.py:18-31."""
    P = polar_factor(X)
    D, N, _, V = P.shape
    det = np.linalg.det(P.transpose(0, 3, 1, 2))  # (D, V)
    root = np.power(det.astype(np.complex128), 1.0 / N)
    return P / root[:, None, None, :]
