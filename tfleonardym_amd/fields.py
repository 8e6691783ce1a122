"""Seeded synthetic inputs: gauge fields and sources (NumPy, complex128, CPU).

Every generator takes an explicit seed and uses ``numpy.random.Generator(PCG64)``
so the device path and any CPU check consume bit-identical arrays.  Layouts
follow the reference: links (D, Nc, Nc, V) (lattice.py:7, hmc.py:20-24),
fundamental spinors (Ns, Nc, V) (hmc.py:199-202), adjoint spinors
(Ns, Nc, Nc, V) as X = sum_a x^a T^a (hmc.py:203-210).
"""
import numpy as np

_SQRT3 = 1.0 / np.sqrt(3.0)


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def volume(geometry):
    return int(np.prod([int(L) for L in geometry]))


def fundamental_generators(nc):
    """T^a = sigma^a/2 (SU(2)) and lambda^a/2 (SU(3)) (lie_generators.py:3-16)."""
    if nc == 2:
        g = np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]]])
    elif nc == 3:
        g = np.array([[[0, 1, 0], [1, 0, 0], [0, 0, 0]],
                      [[0, 0, 1], [0, 0, 0], [1, 0, 0]],
                      [[0, 0, 0], [0, 0, 1], [0, 1, 0]],
                      [[0, -1j, 0], [1j, 0, 0], [0, 0, 0]],
                      [[0, 0, -1j], [0, 0, 0], [1j, 0, 0]],
                      [[0, 0, 0], [0, 0, -1j], [0, 1j, 0]],
                      [[1, 0, 0], [0, -1, 0], [0, 0, 0]],
                      [[_SQRT3, 0, 0], [0, _SQRT3, 0], [0, 0, -2 * _SQRT3]]])
    else:
        raise RuntimeError(f"Lie generators only exist for Nc in (2, 3), got {nc}")
    return (g / 2.0).astype(np.complex128)


def haar_sun(rng, count, nc):
    """Haar-random SU(N) matrices (count, nc, nc): QR of a complex Ginibre
    matrix with the diagonal phase fix (Mezzadri 2007), then det^(-1/N)."""
    z = (rng.standard_normal((count, nc, nc)) + 1j * rng.standard_normal((count, nc, nc))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    q = q * (det ** (-1.0 / nc))[:, None, None]
    return q


def near_identity_sun(rng, count, nc, eps):
    """U = exp(i eps H), H = (G + G^dag)/2 with its trace removed, G complex
    Gaussian with unit-variance real and imaginary parts.  Exact via the
    Hermitian eigendecomposition H = W diag(l) W^dag."""
    g = rng.standard_normal((count, nc, nc)) + 1j * rng.standard_normal((count, nc, nc))
    h = 0.5 * (g + np.conj(np.swapaxes(g, 1, 2)))
    tr = np.trace(h, axis1=1, axis2=2).real / nc
    h = h - tr[:, None, None] * np.eye(nc)[None]
    lam, w = np.linalg.eigh(h)
    ph = np.exp(1j * eps * lam)
    return np.einsum("cij,cj,ckj->cik", w, ph, np.conj(w))


def gauge_field(geometry, nc, seed, kind="haar", eps=0.25):
    """(D, Nc, Nc, V) complex128 links.  kind: "haar", "near_identity" or "unit"."""
    D = len(geometry)
    V = volume(geometry)
    if kind == "unit":
        u = np.zeros((D, nc, nc, V), dtype=np.complex128)
        for a in range(nc):
            u[:, a, a, :] = 1.0
        return u
    rng = _rng(seed)
    if kind == "haar":
        m = haar_sun(rng, D * V, nc)
    elif kind == "near_identity":
        m = near_identity_sun(rng, D * V, nc, eps)
    else:
        raise ValueError(kind)
    # (D*V, nc, nc) with site fastest inside each direction -> (D, nc, nc, V)
    return np.ascontiguousarray(m.reshape(D, V, nc, nc).transpose(0, 2, 3, 1))


def gaussian_spinor(geometry, ns, nc, seed, representation="fundamental", stddev=1.0):
    """Gaussian complex source, re and im ~ N(0, stddev^2) (hmc.py:196-210)."""
    rng = _rng(seed)
    V = volume(geometry)
    if representation == "fundamental":
        return (rng.normal(0, stddev, (ns, nc, V)) + 1j * rng.normal(0, stddev, (ns, nc, V)))
    x = rng.normal(0, stddev, (ns, nc * nc - 1, V)) + 1j * rng.normal(0, stddev, (ns, nc * nc - 1, V))
    return np.einsum("mic,ijk->mjkc", x, fundamental_generators(nc))


def z2_noise(geometry, ns, nc, seed, representation="fundamental"):
    """+-1 real noise per component (fermion_measurements.py:11-23)."""
    rng = _rng(seed)
    V = volume(geometry)
    if representation == "fundamental":
        return (rng.integers(0, 2, (ns, nc, V)) * 2.0 - 1.0).astype(np.complex128)
    x = (rng.integers(0, 2, (ns, nc * nc - 1, V)) * 2.0 - 1.0).astype(np.complex128)
    return np.einsum("mic,ijk->mjkc", x, fundamental_generators(nc))


def point_source(geometry, ns, nc, site, color, spin, representation="fundamental"):
    """Unit source at (spin, color, site) (fermion_measurements.py:26-38)."""
    V = volume(geometry)
    if representation == "fundamental":
        v = np.zeros((ns, nc, V), dtype=np.complex128)
        v[spin, color, site] = 1.0
        return v
    x = np.zeros((ns, nc * nc - 1, V), dtype=np.complex128)
    x[spin, color, site] = 1.0
    return np.einsum("mic,ijk->mjkc", x, fundamental_generators(nc))


def log_shifts(lo, hi, n):
    """n shifts log-spaced in [lo, hi], returned largest first so the base
    (smallest) shift is last, the order the reference's solver expects."""
    return [float(s) for s in np.geomspace(lo, hi, n)[::-1]]
