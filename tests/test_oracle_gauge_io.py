"""Pins the oracle's reader of the "U* stored" SU(3) gauge files
(lattice.py:10-70): round trip of exact SU(3) links, the Gram-Schmidt
reconstruction of skewed columns, and the float decoding against the
standard library's XDR unpacker."""
import struct
import warnings

import numpy as np
import pytest

from oracle import gauge_io as og
from tfleonardym_amd import fields


def _su3_checks(U):
    D, _, _, V = U.shape
    M = U.transpose(0, 3, 1, 2).reshape(-1, 3, 3)
    eye = np.eye(3)
    assert np.abs(np.einsum("nij,nkj->nik", M, np.conj(M)) - eye).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14


@pytest.mark.parametrize("geometry", [(2, 2, 2, 4), (3, 4)])
def test_round_trip_exact_su3(tmp_path, geometry):
    U = fields.gauge_field(geometry, 3, seed=5, kind="haar")
    path = tmp_path / "cfg.xdr"
    og.write_ustar(path, U, geometry, params=(5.6, 0.15, 1.0))
    got, dims, params = og.read_config(path, geometry)
    assert dims == list(geometry) and params == [5.6, 0.15, 1.0]
    assert np.abs(got - U).max() < 3e-7  # float32 storage
    _su3_checks(got)


def test_gram_schmidt_of_skewed_columns(tmp_path):
    geometry = (2, 4)
    V = 8
    rng = np.random.default_rng(2)
    A = rng.standard_normal((2, 3, 3, V)) + 1j * rng.standard_normal((2, 3, 3, V))
    path = tmp_path / "skew.xdr"
    og.write_ustar(path, A, geometry)
    got, _, _ = og.read_config(path, geometry)
    _su3_checks(got)
    a32 = A.astype(np.complex64).astype(np.complex128)
    for mu in range(2):
        for s in range(V):
            a0, a1 = a32[mu, :, 0, s], a32[mu, :, 1, s]
            g0, g1 = got[mu, :, 0, s], got[mu, :, 1, s]
            # column 0 is a0 / |a0|; column 1 the normalised part of a1 orthogonal to it
            assert np.allclose(g0, a0 / np.linalg.norm(a0), atol=1e-14)
            r = a1 - np.vdot(g0, a1) * g0
            assert np.allclose(g1, r / np.linalg.norm(r), atol=1e-13)


def test_float_decoding_matches_xdrlib(tmp_path):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", DeprecationWarning)
        xdrlib = pytest.importorskip("xdrlib")
    vals = np.array([1.5, -2.25, 3.1e-3, 7.0e5], dtype=np.float32)
    data = vals.astype(">f4").tobytes() + struct.pack(">2i", 4, 6)
    p = xdrlib.Unpacker(data)
    ref = p.unpack_farray(4, p.unpack_float)
    assert list(og.read_words(data, 4)) == ref
    assert [p.unpack_int() for _ in range(2)] == [4, 6]


def test_word_order_against_an_independent_writer(tmp_path):
    """The U* file layout pinned by a writer that shares no code with the
    oracle's: explicit loops over (site, mu, row, col < 2) emitting the
    big-endian floats re(conj U[mu][row][col]), im(...) one by one with
    struct.pack, as lattice.py:10-70 reads them (row index before column,
    the conjugate stored).  A row/column transposition or a missing
    conjugation in the oracle's reader fails here."""
    geometry = (2, 3)
    V = 6
    U = fields.gauge_field(geometry, 3, seed=21, kind="haar")
    path = tmp_path / "loop.xdr"
    with open(path, "wb") as f:
        for site in range(V):
            for mu in range(2):
                for row in range(3):
                    for col in range(2):
                        z = np.conj(U[mu, row, col, site])
                        f.write(struct.pack(">f", float(z.real)))
                        f.write(struct.pack(">f", float(z.imag)))
        f.write(struct.pack(">2i", *geometry))
        f.write(struct.pack(">3d", 1.0, 2.0, 3.0))
    got, dims, params = og.read_config(path, geometry)
    assert dims == list(geometry) and params == [1.0, 2.0, 3.0]
    assert np.abs(got - U).max() < 3e-7
    # a transposed file (column index before row) is a different field
    path2 = tmp_path / "transposed.xdr"
    with open(path2, "wb") as f:
        for site in range(V):
            for mu in range(2):
                for col in range(2):
                    for row in range(3):
                        z = np.conj(U[mu, row, col, site])
                        f.write(struct.pack(">f", float(z.real)))
                        f.write(struct.pack(">f", float(z.imag)))
        f.write(struct.pack(">2i", *geometry))
        f.write(struct.pack(">3d", 1.0, 2.0, 3.0))
    bad, _, _ = og.read_config(path2, geometry)
    assert np.abs(bad - U).max() > 1e-2
