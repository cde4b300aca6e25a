"""ctypes front end of the CPU restatement (oracle/eppx_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared against and
the CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module.

Besides loading the C restatement it restates, in numpy, the host-side pieces
that produce the inputs in the reference's own layout:

* ``named_index``        -- NamedIndexTable offsets, src/Class/ParticleSystem/NamedIndex.jl:50-91
* ``rectangle_create``   -- Geometry.create(gap, ::Rectangle), src/Geometry/Geometry2D.jl:71-101
* ``dam_break_case``     -- example/collapse_dry/same/casegen.jl:17-160
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libeppx_oracle.so")

KERNEL_CUBIC, KERNEL_WENDLAND_C2 = 0, 1
EOS_LINEAR, EOS_TAIT = 0, 1
VISC_CLASSIC, VISC_ARTIFICIAL = 0, 1
CRIT_DEFAULT, CRIT_SYMMETRY = 0, 1
ERR_CELL_OVERFLOW, ERR_NEIGHBOUR_OVERFLOW = 1, 2


SOURCES = ("eppx_oracle.c", "eppx_oracle3d.c", "eppx_oracle.h")


def _newest_source() -> float:
    return max(os.path.getmtime(os.path.join(HERE, f)) for f in SOURCES)


def build(force: bool = False) -> str:
    """Compile the restatement (2D and 3D) with its Makefile (gcc; seconds)."""
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < _newest_source():
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class Domain2D(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("gap", "gap_square")] + \
               [(n, C.c_int32) for n in ("n_x", "n_y", "n")] + \
               [(n, C.c_float) for n in ("first_x", "last_x", "first_y", "last_y", "span_x", "span_y",
                                         "gap_x", "gap_y", "gap_x_inv", "gap_y_inv")]


LAYOUT_FIELDS = ("n_capacity", "n_int", "n_float", "max_neighbour_number", "Tag", "IsMovable", "nCount", "nIndex",
                 "PositionVec", "VelocityVec", "dVelocityVec", "Mass", "Volume", "Density", "dDensity", "Pressure",
                 "H", "Gap", "SumWeight", "SumWeightedDensity", "nRVec", "nR", "nW", "nDW", "nHInv",
                 "SumWeightedPressure")


class Layout(C.Structure):
    _fields_ = [(n, C.c_int32) for n in LAYOUT_FIELDS]


class Physics(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("fluid_tag", "wall_tag", "kernel", "eos", "visc")] + \
               [(n, C.c_float) for n in ("dt", "rho0", "p0", "c0", "tait_b", "gx", "gy", "mu", "alpha", "beta",
                                         "av_c", "coefficient")]


class NeighbourSystem(C.Structure):
    _fields_ = [("domain", Domain2D), ("n_cells", C.c_int32), ("max_contained", C.c_int32),
                ("contained_count", C.POINTER(C.c_int32)), ("contained_list", C.POINTER(C.c_int32)),
                ("neighbour_cell_count", C.POINTER(C.c_int32)), ("neighbour_cell_list", C.POINTER(C.c_int32)),
                ("max_tag", C.c_int32), ("adjacency", C.POINTER(C.c_int32))]


_lib = None
_variants = {}


def _bind(L):
    P = C.POINTER
    i32p = P(C.c_int32)
    f32p = P(C.c_float)
    L.or_domain2d_init.argtypes = [C.c_double] * 5 + [P(Domain2D)]
    L.or_index_cartesian_from_position.argtypes = [P(Domain2D), C.c_float, C.c_float, i32p, i32p]
    L.or_index_linear_from_position.argtypes = [P(Domain2D), C.c_float, C.c_float]
    L.or_index_linear_from_position.restype = C.c_int32
    L.or_inside.argtypes = [P(Domain2D), C.c_float, C.c_float]
    L.or_inside.restype = C.c_int32
    L.or_ns_create.argtypes = [P(Domain2D), C.c_int32, i32p, C.c_int32, P(NeighbourSystem)]
    L.or_ns_create.restype = C.c_int32
    L.or_ns_destroy.argtypes = [P(NeighbourSystem)]
    for fn in ("or_kernel_value", "or_kernel_gradient"):
        getattr(L, fn).argtypes = [C.c_int32, C.c_float, C.c_float]
        getattr(L, fn).restype = C.c_float
    L.or_kernel_value0.argtypes = [C.c_int32, C.c_float]
    L.or_kernel_value0.restype = C.c_float
    for fn in ("or_harmonic", "or_invharmonic"):
        getattr(L, fn).argtypes = [C.c_float, C.c_float]
        getattr(L, fn).restype = C.c_float
    L.or_search.argtypes = [P(Layout), C.c_int32, i32p, i32p, i32p, f32p, P(NeighbourSystem), C.c_int32,
                            C.c_int32, P(Physics), C.c_int32, C.c_int32]
    L.or_search.restype = C.c_int32
    L.or_selfaction.argtypes = [P(Layout), C.c_int32, i32p, i32p, f32p, C.c_int32, P(Physics), C.c_int32,
                                C.c_float, C.c_int32]
    L.or_interaction.argtypes = L.or_selfaction.argtypes
    L.or_run_steps.argtypes = [P(Layout), C.c_int32, i32p, i32p, i32p, f32p, P(NeighbourSystem),
                               P(Physics), C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(C.c_int64)]
    L.or_run_steps.restype = C.c_int32
    L.or_set_pair_order.argtypes = [C.c_int32]
    L.or_pair_term_magnitude.argtypes = [P(Layout), C.c_int32, i32p, i32p, f32p, P(Physics), C.c_int32,
                                         P(C.c_double)]
    return L


def lib():
    global _lib
    if _lib is None:
        _lib = _bind(C.CDLL(build()))
    return _lib


def mutant(k: int):
    """Restatement with planted mistake OR_MUTATION=k (for tests that show a check catches it)."""
    if k not in _variants:
        path = os.path.join(HERE, "build", f"libeppx_oracle_m{k}.so")
        if not os.path.exists(path) or os.path.getmtime(path) < _newest_source():
            subprocess.run(["make", "-s", "-C", HERE, f"build/libeppx_oracle_m{k}.so"], check=True)
        _variants[k] = _bind(C.CDLL(path))
    return _variants[k]


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


# ------------------------------------------------------------------ layout

def named_index(named_tuple: dict) -> tuple[dict, int]:
    """1-based start column of each name and the total width (NamedIndex.jl:50-91)."""
    index, start = {}, 1
    for name, width in named_tuple.items():
        index[name] = start
        start += int(width)
    return index, start - 1


def make_layout(int_nt: dict, float_nt: dict, n_capacity: int, basic_map: dict | None = None,
                max_neighbour_number: int | None = None) -> Layout:
    """Reference column table -> 0-based Layout.  ``basic_map`` renames basic
    fields the way basic_index_map_dict does (BasicIndex.jl:20-44)."""
    ii, n_int = named_index(int_nt)
    fi, n_float = named_index(float_nt)
    basic_map = basic_map or {}
    L = Layout()
    L.n_capacity, L.n_int, L.n_float = n_capacity, n_int, n_float
    for f in LAYOUT_FIELDS[4:]:
        name = basic_map.get(f, f)
        if name in ii:
            setattr(L, f, ii[name] - 1)
        elif name in fi:
            setattr(L, f, fi[name] - 1)
        else:
            setattr(L, f, -1)
    if max_neighbour_number is None:
        max_neighbour_number = int(int_nt.get(basic_map.get("nIndex", "nIndex"), 0))
    L.max_neighbour_number = max_neighbour_number
    return L


def domain2d(gap, fx, fy, lx, ly) -> Domain2D:
    d = Domain2D()
    lib().or_domain2d_init(float(gap), float(fx), float(fy), float(lx), float(ly), C.byref(d))
    return d


def physics(**kw) -> Physics:
    p = Physics()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------- geometry

def rectangle_create(gap: float, x1: float, y1: float, x2: float, y2: float):
    """Geometry.create(gap, Rectangle{Float32}(x1, y1, x2, y2)), Geometry2D.jl:13-16 and :71-101.

    Rectangle stores FT(x1) and FT(x2 - x1) (difference taken in Float64);
    n_x = round(Int, span / gap) with round-half-even; dx = span / n_x in
    Float32; x = FT(Float64(first_x) + (mod(i, n_x) - 0.5) * Float64(dx)).
    """
    f32 = np.float32
    first_x, first_y = f32(x1), f32(y1)
    span_x, span_y = f32(x2 - x1), f32(y2 - y1)
    n_x = int(np.rint(np.float64(span_x) / gap))
    n_y = int(np.rint(np.float64(span_y) / gap))
    n = n_x * n_y
    dx = f32(span_x / f32(n_x))
    dy = f32(span_y / f32(n_y))
    volume = f32(dx * dy)
    g = f32(np.sqrt(f32(dx * dy)))
    i = np.arange(1, n + 1, dtype=np.int64)
    i_x = np.mod(i, n_x)
    i_y = (i - 1) // n_x + 1
    x = (np.float64(first_x) + (i_x - 0.5) * np.float64(dx)).astype(np.float32)
    y = (np.float64(first_y) + (i_y - 0.5) * np.float64(dy)).astype(np.float32)
    return np.stack([x, y], axis=1), np.full(n, volume, np.float32), np.full(n, g, np.float32)


def dam_break_layout(M: int):
    """Column tables of casegen.jl:48-70 (2D)."""
    int_nt = {"Tag": 1, "IsMovable": 1, "nCount": 1, "nIndex": M}
    float_nt = {"PositionVec": 2, "VelocityVec": 2, "dVelocityVec": 2, "AccelerationVec": 2, "Mass": 1,
                "Volume": 1, "Density": 1, "dDensity": 1, "Pressure": 1, "Gap": 1, "H": 1, "SumWeight": 1,
                "SumWeightedDensity": 1, "SumWeightedPressure": 1, "nW": M, "nDW": M, "nHInv": M, "nRVec": 2 * M,
                "nR": M}
    return int_nt, float_nt


@dataclass
class Case:
    IP: np.ndarray          # (cap, n_int) int32, Fortran order
    FP: np.ndarray          # (cap, n_float) float32, Fortran order
    is_alive: np.ndarray
    cell_index: np.ndarray
    n_particles: int
    layout: Layout
    domain_args: tuple      # (gap, fx, fy, lx, ly) in Float64
    int_nt: dict
    float_nt: dict


def dam_break_case(dx: float, h_factor: float, M: int, fluid_width=1.0, fluid_height=2.0, box_width=4.0,
                   box_height=3.0, rho0=1000.0, p0=0.0, wall_layers=3, domain_gap=None) -> Case:
    """casegen.jl:17-160 with particle_gap = dx, H = Gap * h_factor and a
    domain margin / cell size of 2h (casegen uses 3dx = 2h for h = 1.5dx)."""
    x0 = y0 = 0.0
    wall_width = wall_layers * dx
    if domain_gap is None:
        domain_gap = 2.0 * h_factor * dx
    rects = [  # merge order fluid, left, right, bottom (casegen.jl:145)
        ((x0, y0, x0 + fluid_width, y0 + fluid_height), 1, 1),
        ((x0 - wall_width, y0, x0, y0 + box_height), 2, 0),
        ((x0 + box_width, y0, x0 + box_width + wall_width, y0 + box_height), 2, 0),
        ((x0 - wall_width, y0 - wall_width, x0 + box_width + wall_width, y0), 2, 0),
    ]
    parts = [(rectangle_create(dx, *r), tag, mov) for r, tag, mov in rects]
    n = sum(p[0][0].shape[0] for p in parts)
    int_nt, float_nt = dam_break_layout(M)
    L = make_layout(int_nt, float_nt, n)
    IP = np.zeros((n, L.n_int), np.int32, order="F")
    FP = np.zeros((n, L.n_float), np.float32, order="F")
    o = 0
    for (pos, vol, gap), tag, mov in parts:
        k = pos.shape[0]
        s = slice(o, o + k)
        FP[s, L.PositionVec] = pos[:, 0]
        FP[s, L.PositionVec + 1] = pos[:, 1]
        FP[s, L.Volume] = vol
        FP[s, L.Gap] = gap
        IP[s, L.Tag] = tag
        IP[s, L.IsMovable] = mov
        FP[s, L.Density] = np.float32(rho0)
        FP[s, L.Mass] = FP[s, L.Density] * FP[s, L.Volume]
        FP[s, L.H] = (FP[s, L.Gap].astype(np.float64) * h_factor).astype(np.float32)
        FP[s, L.Pressure] = np.float32(p0)
        o += k
    dom = (domain_gap, x0 - domain_gap, y0 - domain_gap, x0 + box_width + domain_gap, y0 + box_height + domain_gap)
    return Case(IP, FP, np.ones(n, np.int32), np.zeros(n, np.int32), n, L, dom, int_nt, float_nt)


# ------------------------------------------------------------------ system

class System:
    """Reference-layout particle + neighbour system driven by the restatement."""

    def __init__(self, IP, FP, is_alive, cell_index, n_particles, layout: Layout, domain: Domain2D,
                 adjacency: np.ndarray, max_contained: int, variant=None):
        self.lib = variant if variant is not None else lib()
        self.IP = np.asfortranarray(IP, dtype=np.int32)
        self.FP = np.asfortranarray(FP, dtype=np.float32)
        self.is_alive = np.ascontiguousarray(is_alive, dtype=np.int32)
        self.cell_index = np.ascontiguousarray(cell_index, dtype=np.int32)
        self.n = int(n_particles)
        self.L = layout
        self.L.n_capacity = self.IP.shape[0]
        self.domain = domain
        adj = np.asfortranarray(adjacency, dtype=np.int32)
        self.ns = NeighbourSystem()
        rc = self.lib.or_ns_create(C.byref(domain), int(max_contained), _ptr(adj, C.c_int32), adj.shape[0],
                                C.byref(self.ns))
        if rc != 0:
            raise MemoryError("or_ns_create failed")

    def __del__(self):
        if getattr(self, "ns", None) is not None and getattr(self, "lib", None) is not None:
            self.lib.or_ns_destroy(C.byref(self.ns))
            self.ns = None

    def _args(self):
        return (C.byref(self.L), self.n, _ptr(self.is_alive, C.c_int32))

    def search(self, criterion=CRIT_SYMMETRY, action=0, phys: Physics | None = None, serial=True, threads=0):
        phys = phys or Physics()
        return self.lib.or_search(C.byref(self.L), self.n, _ptr(self.is_alive, C.c_int32),
                               _ptr(self.cell_index, C.c_int32), _ptr(self.IP, C.c_int32), _ptr(self.FP, C.c_float),
                               C.byref(self.ns), criterion, action, C.byref(phys), int(serial), threads)

    def selfaction(self, action, phys: Physics | None = None, column=0, value=0.0, threads=0):
        self.lib.or_selfaction(*self._args(), _ptr(self.IP, C.c_int32), _ptr(self.FP, C.c_float), action,
                            C.byref(phys or Physics()), column, value, threads)

    def interaction(self, action, phys: Physics | None = None, column=0, value=0.0, threads=0):
        self.lib.or_interaction(*self._args(), _ptr(self.IP, C.c_int32), _ptr(self.FP, C.c_float), action,
                             C.byref(phys or Physics()), column, value, threads)

    def run_steps(self, n_steps, phys: Physics, filter_interval=0, serial=True, threads=0):
        pairs = np.zeros(max(n_steps, 1), np.int64)
        err = self.lib.or_run_steps(C.byref(self.L), self.n, _ptr(self.is_alive, C.c_int32),
                                 _ptr(self.cell_index, C.c_int32), _ptr(self.IP, C.c_int32),
                                 _ptr(self.FP, C.c_float), C.byref(self.ns), C.byref(phys), n_steps,
                                 filter_interval, int(serial), threads, _ptr(pairs, C.c_int64))
        return err, pairs[:n_steps]

    def term_magnitude(self, what, phys: Physics):
        """Sum of |term| over each particle's neighbours: what 0 the continuity terms (n,), 1 the momentum
        terms (n, 2)."""
        out = np.zeros((self.n, 2) if what else self.n, np.float64)
        self.lib.or_pair_term_magnitude(*self._args(), _ptr(self.IP, C.c_int32), _ptr(self.FP, C.c_float),
                                        C.byref(phys), int(what), _ptr(out, C.c_double))
        return out

    def set_pair_order(self, reversed_):
        """Summation order of this library's per-neighbour sums (0 list order, 1 reversed)."""
        self.lib.or_set_pair_order(int(reversed_))

    # -------------------------------------------------------------- views
    def count(self):
        return self.IP[: self.n, self.L.nCount].copy()

    def neighbours(self, i):
        """1-based neighbour indices of 0-based particle i, in list order."""
        c = self.IP[i, self.L.nCount]
        return self.IP[i, self.L.nIndex: self.L.nIndex + c].copy()

    def pair_csr(self):
        """(offsets, sorted 0-based neighbour ids) over particles 0..n-1 (dead rows empty)."""
        cnt = np.where(self.is_alive[: self.n] == 1, self.IP[: self.n, self.L.nCount], 0).astype(np.int64)
        off = np.zeros(self.n + 1, np.int64)
        np.cumsum(cnt, out=off[1:])
        M = self.L.max_neighbour_number
        lists = self.IP[: self.n, self.L.nIndex: self.L.nIndex + M] - 1
        mask = np.arange(M)[None, :] < cnt[:, None]
        js = np.where(mask, lists, np.iinfo(np.int32).max)
        js.sort(axis=1)
        return off, js[mask.sum(axis=1)[:, None] > np.arange(M)[None, :]].astype(np.int32)

    def column(self, name):
        return self.FP[: self.n, getattr(self.L, name)].copy()
