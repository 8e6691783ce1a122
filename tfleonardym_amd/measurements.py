"""Measurement record of the reference (measurements.py:7-8): the fermion
measurements return it.  The gauge observables of that module are outside
the fermion-matrix path and are not part of this package."""
import collections

Measurement = collections.namedtuple("Measurement", ["geometry", "colors", "value", "name", "smearing_levels"])
Measurement.__new__.__defaults__ = (None, None, None, None, 0)
