"""Geometry: host lattice generators used to build cases (src/Geometry/Geometry2D.jl).

Only the rectangle is on the dam-break path.  The arithmetic keeps the
reference's mixed precision: Rectangle{Float32} stores FT(x1) and FT(x2 - x1)
(difference in Float64); count / create use round-half-even of
Float64(span) / gap; dx = span / n_x in Float32; positions are
FT(Float64(first) + (mod(i, n_x) - 0.5) * Float64(dx)).
"""
from __future__ import annotations

import numpy as np


class Rectangle:
    """Rectangle{Float32}(x1, y1, x2, y2), Geometry2D.jl:10-25."""

    def __init__(self, x1, y1, x2, y2):
        assert x1 < x2, "x1 must be less than x2"
        assert y1 < y2, "y1 must be less than y2"
        self.first_x, self.first_y = np.float32(x1), np.float32(y1)
        self.last_x, self.last_y = np.float32(x2), np.float32(y2)
        self.span_x, self.span_y = np.float32(x2 - x1), np.float32(y2 - y1)


def _n(span, gap):
    return int(np.rint(np.float64(span) / gap))  # Int(round(x)): ties to even


def count(gap: float, rect: Rectangle) -> int:
    """Geometry.count, Geometry2D.jl:58-62."""
    return _n(rect.span_x, gap) * _n(rect.span_y, gap)


def inside(x, y, rect: Rectangle) -> bool:
    """Geometry.inside, Geometry2D.jl:64-69."""
    return rect.first_x <= x <= rect.last_x and rect.first_y <= y <= rect.last_y


def create(gap: float, rect: Rectangle):
    """Geometry.create, Geometry2D.jl:71-101: (positions (n, 2), volumes, gaps), Float32.

    Row-major lattice with i_x = mod(i, n_x): the last particle of each row
    sits at first_x - dx / 2 (a quirk the dam-break case inherits).
    """
    n_x, n_y = _n(rect.span_x, gap), _n(rect.span_y, gap)
    dx = np.float32(rect.span_x / np.float32(n_x))
    dy = np.float32(rect.span_y / np.float32(n_y))
    vol = np.float32(dx * dy)
    g = np.float32(np.sqrt(np.float32(dx * dy)))
    i = np.arange(1, n_x * n_y + 1, dtype=np.int64)
    i_x = np.mod(i, n_x)
    i_y = (i - 1) // n_x + 1
    x = (np.float64(rect.first_x) + (i_x - 0.5) * np.float64(dx)).astype(np.float32)
    y = (np.float64(rect.first_y) + (i_y - 0.5) * np.float64(dy)).astype(np.float32)
    n = x.shape[0]
    return np.stack([x, y], axis=1), np.full(n, vol, np.float32), np.full(n, g, np.float32)


class Cuboid:
    """Cuboid{Float32}(x1, y1, z1, x2, y2, z2) (Geometry3D.jl:10-30): first corner and spans, as Rectangle."""

    def __init__(self, x1, y1, z1, x2, y2, z2):
        assert x1 < x2 and y1 < y2 and z1 < z2, "the first corner must lie below the last"
        self.first_x, self.first_y, self.first_z = np.float32(x1), np.float32(y1), np.float32(z1)
        self.span_x, self.span_y, self.span_z = np.float32(x2 - x1), np.float32(y2 - y1), np.float32(z2 - z1)


def count3(gap: float, cub: Cuboid) -> int:
    """Geometry.count for a Cuboid (Geometry3D.jl:74-79): the product of round(span / gap) per axis."""
    return _n(cub.span_x, gap) * _n(cub.span_y, gap) * _n(cub.span_z, gap)


def create3(gap: float, cub: Cuboid):
    """A cuboid lattice: (positions (n, 3), volumes, gaps), Float32.

    The reference's Cuboid create stops at the counts (Geometry3D.jl:89-94
    "TODO"), so this lattice extends the rectangle's arithmetic per axis
    (dx = span_x / n_x in Float32, x = FT(Float64(first_x) + (i_x - 0.5)
    Float64(dx)), volume dx dy dz, gap the cube root of the volume) with
    i_x = mod1(i, n_x), i_y = mod1(cld(i, n_x), n_y), i_z = cld(i, n_x n_y):
    x fastest, and no wrapped last column.
    """
    n_x, n_y, n_z = _n(cub.span_x, gap), _n(cub.span_y, gap), _n(cub.span_z, gap)
    dx = np.float32(cub.span_x / np.float32(n_x))
    dy = np.float32(cub.span_y / np.float32(n_y))
    dz = np.float32(cub.span_z / np.float32(n_z))
    vol = np.float32(np.float32(dx * dy) * dz)
    g = np.float32(np.float64(vol) ** (1.0 / 3.0))
    i = np.arange(n_x * n_y * n_z, dtype=np.int64)
    i_x = i % n_x + 1
    i_y = (i // n_x) % n_y + 1
    i_z = i // (n_x * n_y) + 1
    x = (np.float64(cub.first_x) + (i_x - 0.5) * np.float64(dx)).astype(np.float32)
    y = (np.float64(cub.first_y) + (i_y - 0.5) * np.float64(dy)).astype(np.float32)
    z = (np.float64(cub.first_z) + (i_z - 0.5) * np.float64(dz)).astype(np.float32)
    n = x.shape[0]
    return np.stack([x, y, z], axis=1), np.full(n, vol, np.float32), np.full(n, g, np.float32)
