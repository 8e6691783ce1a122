"""CPU check of the kernels' per-site Wilson arithmetic (wilson_site.h compiled
for the host, looped over the lattice) against the oracle: site indexing,
gamma tables, colour transport and the in-kernel boundary factor, for every
(ndim, Nc, representation, hermitian, boundary, mode) the kernels build."""
import ctypes
import os

import numpy as np
import pytest

from oracle import wilson as ow
from tfleonardym_amd import fields

LIB = os.path.join(os.path.dirname(__file__), "..", "tfleonardym_amd", "liblym_cpucheck.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from tfleonardym_amd import build
        build.build_cpu_check()
    lib = ctypes.CDLL(LIB)
    lib.lymcheck_wilson_apply.restype = ctypes.c_int
    return lib


def run(lib, geometry, nc, adj, herm, kappa, bc, proj, links, links_r, psi):
    ndim = len(geometry)
    dims = (ctypes.c_int * 4)(*(list(geometry) + [1] * (4 - ndim)))
    out = np.zeros_like(psi)
    links = np.ascontiguousarray(links)
    lr = None if links_r is None else np.ascontiguousarray(links_r)
    rc = lib.lymcheck_wilson_apply(ndim, nc, adj, herm, dims, ctypes.c_double(kappa), ctypes.c_double(bc.real),
                                   ctypes.c_double(bc.imag), proj, links.ctypes.data_as(ctypes.c_void_p),
                                   None if lr is None else lr.ctypes.data_as(ctypes.c_void_p),
                                   np.ascontiguousarray(psi).ctypes.data_as(ctypes.c_void_p),
                                   out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    return out


CASES = []
for geometry in [(4, 6), (3, 4), (2, 4, 2, 4), (4, 2, 3, 4)]:
    for nc in (2, 3):
        for adj in (0, 1):
            CASES.append((geometry, nc, adj))


@pytest.mark.parametrize("geometry,nc,adj", CASES)
@pytest.mark.parametrize("herm", [0, 1])
@pytest.mark.parametrize("bc", [None, -1.0, np.exp(0.3j)])
@pytest.mark.parametrize("proj", [1, 0])
def test_site_arithmetic_matches_oracle(lib, geometry, nc, adj, herm, bc, proj):
    D = len(geometry)
    ns = 2 if D == 2 else 4
    rep = "adjoint" if adj else "fundamental"
    U = fields.gauge_field(geometry, nc, seed=13, kind="haar")
    psi = fields.gaussian_spinor(geometry, ns, nc, seed=4, representation=rep)
    kappa = 0.1234
    op = ow.WilsonOracle(U, geometry, kappa, hermitian=bool(herm), representation=rep, bc_factor=bc)
    want = op.apply(psi)
    got = run(lib, geometry, nc, adj, herm, kappa, complex(1.0 if bc is None else bc), proj, U, None, psi)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 1e-14, err


def test_adjoint_custom_provider_two_link_fields(lib):
    """links_right != links: left field carries an arbitrary provider's change."""
    geometry = (4, 4)
    nc = 2
    U = fields.gauge_field(geometry, nc, seed=3, kind="haar")
    X = fields.gaussian_spinor(geometry, 2, nc, seed=1, representation="adjoint")
    Ul = U.copy()
    Ul[0] *= np.exp(0.7j)  # a provider that twists direction 0 everywhere
    want = ow.apply_projected_adjoint(Ul, U, X, 0.11, True, __import__("oracle").lattice.lookup_tables(geometry))
    got = run(lib, geometry, nc, 1, 1, 0.11, 1.0 + 0j, 1, Ul, U, X)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-14


def encode16(psi):
    """16-bit storage of a spinor as the kernels write it: per site the norm
    n = max |component| / 32767 (float32) and int16 = rint(value * 32767 / max)."""
    ns, nc, V = psi.shape
    comp = np.stack([psi.real, psi.imag], axis=-1).reshape(ns * nc, V, 2)
    m = np.abs(comp).max(axis=(0, 2)).astype(np.float32)
    norm = (m * np.float32(1.0 / 32767.0)).astype(np.float32)
    inv = np.where(m > 0, np.float32(32767.0) / np.where(m > 0, m, 1), 0).astype(np.float32)
    ints = np.clip(np.rint(comp.astype(np.float32) * inv[None, :, None]), -32767, 32767).astype(np.int16)
    return ints, norm


def pack16(ints):
    """(NCOMP, V, 2) int16 -> the kernels' packed layout: G complex numbers per
    element (G = 4 if NCOMP % 4 == 0 else 2), component c of site s at
    element (c // G) * V + s, slot c % G."""
    ncomp, V, _ = ints.shape
    g = 4 if ncomp % 4 == 0 else 2
    return np.ascontiguousarray(ints.reshape(ncomp // g, g, V, 2).transpose(0, 2, 1, 3))


@pytest.mark.parametrize("geometry,nc", [((4, 6), 2), ((3, 4), 3), ((2, 4, 2, 4), 2), ((4, 2, 3, 4), 3)])
@pytest.mark.parametrize("herm", [0, 1])
@pytest.mark.parametrize("bc", [None, -1.0])
@pytest.mark.parametrize("proj", [1, 0])
def test_half_storage_hop_matches_oracle(lib, geometry, nc, herm, bc, proj):
    """The 16-bit input path of the hop (neighbour integers converted to float
    and the site norm folded into the accumulating FMAs), complex64 links and
    arithmetic: equal to the oracle applied to the decoded input within
    complex64 rounding."""
    ndim = len(geometry)
    ns = 2 if ndim == 2 else 4
    U = fields.gauge_field(geometry, nc, seed=21, kind="haar")
    psi = fields.gaussian_spinor(geometry, ns, nc, seed=22)
    psi[:, :, 1] *= 1e-3  # sites of very different size keep their own scale
    ints, norm = encode16(psi)
    V = psi.shape[-1]
    dec = (ints[..., 0].astype(np.float64) + 1j * ints[..., 1].astype(np.float64)) * norm[None, :].astype(np.float64)
    dec = dec.reshape(ns, nc, V)
    kappa = 0.1234
    want = ow.WilsonOracle(U, geometry, kappa, hermitian=bool(herm), representation="fundamental",
                           bc_factor=bc).apply(dec)
    lib.lymcheck_wilson_apply_half.restype = ctypes.c_int
    dims = (ctypes.c_int * 4)(*(list(geometry) + [1] * (4 - ndim)))
    links = np.ascontiguousarray(U.astype(np.complex64))
    out = np.zeros(psi.shape, dtype=np.complex64)
    b = 1.0 + 0j if bc is None else complex(bc)
    rc = lib.lymcheck_wilson_apply_half(ndim, nc, herm, dims, ctypes.c_double(kappa), ctypes.c_double(b.real),
                                        ctypes.c_double(b.imag), proj, links.ctypes.data_as(ctypes.c_void_p),
                                        pack16(ints).ctypes.data_as(ctypes.c_void_p),
                                        norm.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    err = np.linalg.norm(out.astype(np.complex128) - want) / np.linalg.norm(want)
    assert err < 2e-6, err


def test_buffer_range_predicate(lib):
    """The host check of buffer arguments (check_buffers in liblym): a buffer
    of `bytes` starting at p must end inside the allocation [base, base+size)
    that contains p.  Exercised on the host build; a short buffer is never
    launched on the GPU."""
    f = lib.cpucheck_range_inside
    f.argtypes = [ctypes.c_uint64] * 4
    f.restype = ctypes.c_int
    base, size = 0x7F0000000000, 4096
    assert f(base, size, base, 4096) == 1          # exactly the allocation
    assert f(base, size, base + 1024, 3072) == 1   # a tail slice
    assert f(base, size, base + 1024, 3073) == 0   # one byte past the end
    assert f(base, size, base, 4097) == 0
    assert f(base, size, base - 8, 16) == 0        # starts before the allocation
    assert f(base, size, base + 4096, 1) == 0      # starts at the end
    assert f(base, size, base + 4096, 0) == 1      # empty range at the end
    assert f(base, size, base + 8, 2**64 - 8) == 0  # no wrap-around
