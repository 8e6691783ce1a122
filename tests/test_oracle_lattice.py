"""Pins the oracle's geometry (lattice.py:86-128, boundary_conditions.py:10-46)
by brute force over coordinates."""
import itertools

import numpy as np
import pytest

from oracle import lattice as olat


@pytest.mark.parametrize("geometry", [(4, 3), (2, 3, 4, 5), (6, 4), (3, 3, 3, 4)])
def test_tables_brute_force(geometry):
    sup, sdn = olat.lookup_tables(geometry)
    D = len(geometry)
    s = 0
    for x in itertools.product(*[range(L) for L in geometry]):
        # last coordinate fastest: itertools.product enumerates in that order
        for mu in range(D):
            xp = list(x)
            xp[mu] = (xp[mu] + 1) % geometry[mu]
            xm = list(x)
            xm[mu] = (xm[mu] - 1) % geometry[mu]
            sp = 0
            sm = 0
            for nu in range(D):
                sp = sp * geometry[nu] + xp[nu]
                sm = sm * geometry[nu] + xm[nu]
            assert sup[mu][s] == sp
            assert sdn[mu][s] == sm
        s += 1
    assert s == np.prod(geometry)


def test_translate_is_value_at_x_plus_mu():
    geometry = (4, 6)
    tables = olat.lookup_tables(geometry)
    c = olat.coordinates(geometry).astype(float)
    # translating the coordinate field by +mu gives coordinate+1 mod L
    for mu in range(2):
        t = olat.translate(c[mu], mu, +1, tables)
        assert np.array_equal(t, (c[mu] + 1) % geometry[mu])
        t = olat.translate(c[mu], mu, -1, tables)
        assert np.array_equal(t, (c[mu] - 1) % geometry[mu])
        # up then down is the identity
        assert np.array_equal(olat.translate(olat.translate(c[1 - mu], mu, +1, tables), mu, -1, tables), c[1 - mu])


def test_antiperiodic_field():
    geometry = (3, 4)
    f = olat.antiperiodic_field(geometry)
    c = olat.coordinates(geometry)
    assert np.all(f[c[-1] == 3] == -1)
    assert np.all(f[c[-1] != 3] == 1)
    assert f.sum() == 12 - 2 * 3
