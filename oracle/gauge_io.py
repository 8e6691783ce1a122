"""Oracle: the "U* stored" SU(3) gauge-file format (test infrastructure only).

read_config restates lattice.py:10-70 with struct in place of xdrlib: the
XDR floats of the two stored columns per link, conjugation, Gram-Schmidt of
the columns in the reference's order, the third column as the conjugated
cross product, and the trailer of extents and parameters.  write_ustar
writes the same format from (D, 3, 3, V) links (test files).

Pinning (tests/test_oracle_gauge_io.py): exact SU(3) links survive
write -> read to float32 rounding and come back exactly special unitary;
skewed, unnormalised columns give the unique SU(3) matrix whose first two
columns are their Gram-Schmidt orthonormalisation; the float decoding
agrees with the standard library's XDR unpacker.
"""
import struct

import numpy as np


def write_ustar(path, links, geometry, params=(0.0, 0.0, 0.0)):
    """(D, 3, 3, V) links -> file: conj of columns 0 and 1 as XDR floats in
    the order (site, mu, row, col, re/im), then the extents and 3 doubles."""
    U = np.asarray(links)
    D, nc, _, V = U.shape
    ustar = np.conj(U[:, :, :2, :]).transpose(3, 0, 1, 2)  # (V, D, row, col) of U*
    words = np.stack([ustar.real, ustar.imag], axis=-1)
    with open(path, "wb") as f:
        f.write(words.astype(">f4").tobytes())
        f.write(struct.pack(f">{D}i", *[int(L) for L in geometry]))
        f.write(struct.pack(">3d", *params))


def read_words(data, n):
    return np.array(struct.unpack(f">{n}f", data[:4 * n]), dtype=np.float64)


def read_config(path, geometry, Nc=3):
    """lattice.py:10-70 -> ((D, 3, 3, V) complex128 links, extents, params)."""
    geometry = tuple(int(L) for L in geometry)
    D = len(geometry)
    V = int(np.prod(geometry))
    n = V * (Nc - 1) * Nc * D * 2
    data = open(path, "rb").read()
    g = read_words(data, n).reshape(V, D, Nc, Nc - 1, 2)
    loaded = g[..., 0] - 1j * g[..., 1]
    out = np.zeros((D, Nc, Nc, V), dtype=np.complex128)
    for mu in range(D):
        c0 = loaded[:, mu, :, 0].T  # (3, V)
        c0 = c0 / np.sqrt(np.sum(np.abs(c0) ** 2, axis=0))
        c1 = loaded[:, mu, :, 1].T
        t = np.sum(c1 * np.conj(c0), axis=0)
        c1 = c1 - t * c0
        c1 = c1 / np.sqrt(np.sum(np.abs(c1) ** 2, axis=0))
        c2 = np.conj(np.stack([c0[1] * c1[2] - c0[2] * c1[1],
                               c0[2] * c1[0] - c0[0] * c1[2],
                               c0[0] * c1[1] - c0[1] * c1[0]]))
        out[mu, :, 0] = c0
        out[mu, :, 1] = c1
        out[mu, :, 2] = c2
    tail = data[4 * n:]
    dims = list(struct.unpack(f">{D}i", tail[:4 * D]))
    params = list(struct.unpack(">3d", tail[4 * D:4 * D + 24]))
    return out, dims, params
