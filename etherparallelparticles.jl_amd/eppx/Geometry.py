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
