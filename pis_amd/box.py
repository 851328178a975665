"""SimulationBox mirror (src/simulation_box.rs:5-69): box matrix h (box vectors as columns,
column-major like nalgebra) and its inverse, produced with the same cofactor formula as
nalgebra 0.34.1's 3x3 try_inverse so the engine sees the bits the reference would."""
from __future__ import annotations

import numpy as np


def _m(a, r, c):
    return a[c * 3 + r]


def box_inverse(h):
    """3x3 inverse of a column-major 9-vector (nalgebra linalg/inverse.rs, 3x3 branch):
    cofactors, determinant by first-row expansion left to right, entries divided by it.
    Plain Python floats so every operation rounds once, like the Rust code."""
    h = [float(x) for x in h]
    m11 = _m(h, 1, 1) * _m(h, 2, 2) - _m(h, 2, 1) * _m(h, 1, 2)
    m12 = _m(h, 1, 0) * _m(h, 2, 2) - _m(h, 2, 0) * _m(h, 1, 2)
    m13 = _m(h, 1, 0) * _m(h, 2, 1) - _m(h, 2, 0) * _m(h, 1, 1)
    det = _m(h, 0, 0) * m11 - _m(h, 0, 1) * m12 + _m(h, 0, 2) * m13
    if det == 0.0:
        raise ValueError("Box matrix should be invertible")
    out = [0.0] * 9

    def put(r, c, v):
        out[c * 3 + r] = v

    put(0, 0, m11 / det)
    put(0, 1, -(_m(h, 0, 1) * _m(h, 2, 2) - _m(h, 2, 1) * _m(h, 0, 2)) / det)
    put(0, 2, (_m(h, 0, 1) * _m(h, 1, 2) - _m(h, 1, 1) * _m(h, 0, 2)) / det)
    put(1, 0, -m12 / det)
    put(1, 1, (_m(h, 0, 0) * _m(h, 2, 2) - _m(h, 2, 0) * _m(h, 0, 2)) / det)
    put(1, 2, -(_m(h, 0, 0) * _m(h, 1, 2) - _m(h, 1, 0) * _m(h, 0, 2)) / det)
    put(2, 0, m13 / det)
    put(2, 1, -(_m(h, 0, 0) * _m(h, 2, 1) - _m(h, 2, 0) * _m(h, 0, 1)) / det)
    put(2, 2, (_m(h, 0, 0) * _m(h, 1, 1) - _m(h, 1, 0) * _m(h, 0, 1)) / det)
    return out


class SimulationBox:
    """h, h_inv and pbc flags; from_lammps_data builds columns a=(ax,0,0), b=(xy,ay,0),
    c=(xz,yz,az) with pbc on everywhere (simulation_box.rs:44-65)."""

    def __init__(self, h, pbc=(True, True, True)):
        self.h = [float(x) for x in h]
        self.h_inv = box_inverse(self.h)
        self.pbc = tuple(bool(p) for p in pbc)

    @classmethod
    def from_lammps_data(cls, xlo, xhi, ylo, yhi, zlo, zhi, xy=0.0, xz=0.0, yz=0.0):
        h = [0.0] * 9
        h[0] = xhi - xlo
        h[3] = xy
        h[4] = yhi - ylo
        h[6] = xz
        h[7] = yz
        h[8] = zhi - zlo
        return cls(h)

    @classmethod
    def cubic(cls, length):
        return cls.from_lammps_data(0.0, length, 0.0, length, 0.0, length)

    @property
    def lengths(self):
        return np.array([self.h[0], self.h[4], self.h[8]])

    def volume(self):
        h = self.h
        m11 = _m(h, 1, 1) * _m(h, 2, 2) - _m(h, 2, 1) * _m(h, 1, 2)
        m12 = _m(h, 1, 0) * _m(h, 2, 2) - _m(h, 2, 0) * _m(h, 1, 2)
        m13 = _m(h, 1, 0) * _m(h, 2, 1) - _m(h, 2, 0) * _m(h, 1, 1)
        return abs(_m(h, 0, 0) * m11 - _m(h, 0, 1) * m12 + _m(h, 0, 2) * m13)
