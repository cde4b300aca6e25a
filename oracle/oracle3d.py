"""ctypes front end of the 3D restatement (oracle/eppx_oracle3d.c).

TEST INFRASTRUCTURE ONLY: the checker the 3D HIP path is compared against
and the CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module.

The reference's 3D layers are stubs (AbstractDomain.jl:119,
NeighbourSystemBase.jl:50-57, NeighbourSearch.jl:49, :229), so the 3D
restatement extends the 2D one axis by axis; it is pinned by closed forms,
invariants and brute force (tests/test_oracle3d.py), not by reference output.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


class Domain3D(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("gap", "gap_square")] + \
               [(n, C.c_int32) for n in ("n_x", "n_y", "n_z", "n")] + \
               [(n, C.c_float) for n in ("first_x", "last_x", "first_y", "last_y", "first_z", "last_z", "span_x",
                                         "span_y", "span_z", "gap_x", "gap_y", "gap_z", "gap_x_inv", "gap_y_inv",
                                         "gap_z_inv")]


class Physics3D(C.Structure):
    _fields_ = [("p", O.Physics), ("gz", C.c_float)]


class NeighbourSystem3D(C.Structure):
    _fields_ = [("domain", Domain3D), ("n_cells", C.c_int32), ("max_contained", C.c_int32),
                ("contained_count", C.POINTER(C.c_int32)), ("contained_list", C.POINTER(C.c_int32)),
                ("neighbour_cell_count", C.POINTER(C.c_int32)), ("neighbour_cell_list", C.POINTER(C.c_int32)),
                ("max_tag", C.c_int32), ("adjacency", C.POINTER(C.c_int32))]


_bound = {}


def _bind(L):
    if id(L) in _bound:
        return L
    P = C.POINTER
    i32p, f32p = P(C.c_int32), P(C.c_float)
    L.or3_domain_init.argtypes = [C.c_double] * 7 + [P(Domain3D)]
    L.or3_index_linear_from_position.argtypes = [P(Domain3D), C.c_float, C.c_float, C.c_float]
    L.or3_index_linear_from_position.restype = C.c_int32
    L.or3_inside.argtypes = [P(Domain3D), C.c_float, C.c_float, C.c_float]
    L.or3_inside.restype = C.c_int32
    L.or3_ns_create.argtypes = [P(Domain3D), C.c_int32, i32p, C.c_int32, P(NeighbourSystem3D)]
    L.or3_ns_create.restype = C.c_int32
    L.or3_ns_destroy.argtypes = [P(NeighbourSystem3D)]
    for fn in ("or3_kernel_value", "or3_kernel_gradient"):
        getattr(L, fn).argtypes = [C.c_int32, C.c_float, C.c_float]
        getattr(L, fn).restype = C.c_float
    L.or3_kernel_value0.argtypes = [C.c_int32, C.c_float]
    L.or3_kernel_value0.restype = C.c_float
    L.or3_search.argtypes = [P(O.Layout), C.c_int32, i32p, i32p, i32p, f32p, P(NeighbourSystem3D), C.c_int32,
                             C.c_int32, P(Physics3D), C.c_int32, C.c_int32]
    L.or3_search.restype = C.c_int32
    L.or3_selfaction.argtypes = [P(O.Layout), C.c_int32, i32p, i32p, f32p, C.c_int32, P(Physics3D), C.c_int32]
    L.or3_interaction.argtypes = L.or3_selfaction.argtypes
    L.or3_run_steps.argtypes = [P(O.Layout), C.c_int32, i32p, i32p, i32p, f32p, P(NeighbourSystem3D),
                                P(Physics3D), C.c_int32, C.c_int32, C.c_int32, P(C.c_int64)]
    L.or3_run_steps.restype = C.c_int32
    L.or3_pair_term_magnitude.argtypes = [P(O.Layout), C.c_int32, i32p, i32p, f32p, P(Physics3D), C.c_int32,
                                          P(C.c_double)]
    _bound[id(L)] = True
    return L


def lib():
    return _bind(O.lib())


def mutant(k: int):
    """The 2D and 3D restatement with planted mistake OR_MUTATION=k."""
    return _bind(O.mutant(k))


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def domain3d(gap, fx, fy, fz, lx, ly, lz, variant=None) -> Domain3D:
    d = Domain3D()
    (variant or lib()).or3_domain_init(float(gap), float(fx), float(fy), float(fz), float(lx), float(ly), float(lz),
                                       C.byref(d))
    return d


def physics3d(gz: float = 0.0, **kw) -> Physics3D:
    return Physics3D(O.physics(**kw), float(np.float32(gz)))


class System3D:
    """Reference-layout 3D particle + neighbour system driven by the restatement."""

    def __init__(self, IP, FP, is_alive, cell_index, n_particles, layout: O.Layout, domain: Domain3D,
                 adjacency: np.ndarray, max_contained: int, variant=None):
        self.lib = _bind(variant) if variant is not None else lib()
        self.IP = np.asfortranarray(IP, dtype=np.int32)
        self.FP = np.asfortranarray(FP, dtype=np.float32)
        self.is_alive = np.ascontiguousarray(is_alive, dtype=np.int32)
        self.cell_index = np.ascontiguousarray(cell_index, dtype=np.int32)
        self.n = int(n_particles)
        self.L = layout
        self.L.n_capacity = self.IP.shape[0]
        self.domain = domain
        adj = np.asfortranarray(adjacency, dtype=np.int32)
        self.ns = NeighbourSystem3D()
        if self.lib.or3_ns_create(C.byref(domain), int(max_contained), _ptr(adj, C.c_int32), adj.shape[0],
                                  C.byref(self.ns)) != 0:
            raise MemoryError("or3_ns_create failed")

    def __del__(self):
        if getattr(self, "ns", None) is not None and getattr(self, "lib", None) is not None:
            self.lib.or3_ns_destroy(C.byref(self.ns))
            self.ns = None

    def _tables(self):
        return (_ptr(self.IP, C.c_int32), _ptr(self.FP, C.c_float))

    def search(self, criterion=O.CRIT_SYMMETRY, action=0, phys: Physics3D | None = None, serial=True, threads=0):
        phys = phys or Physics3D()
        return self.lib.or3_search(C.byref(self.L), self.n, _ptr(self.is_alive, C.c_int32),
                                   _ptr(self.cell_index, C.c_int32), *self._tables(), C.byref(self.ns), criterion,
                                   action, C.byref(phys), int(serial), threads)

    def selfaction(self, action, phys: Physics3D, threads=0):
        self.lib.or3_selfaction(C.byref(self.L), self.n, _ptr(self.is_alive, C.c_int32), *self._tables(), action,
                                C.byref(phys), threads)

    def interaction(self, action, phys: Physics3D, threads=0):
        self.lib.or3_interaction(C.byref(self.L), self.n, _ptr(self.is_alive, C.c_int32), *self._tables(), action,
                                 C.byref(phys), threads)

    def run_steps(self, n_steps, phys: Physics3D, serial=True, threads=0):
        pairs = np.zeros(max(n_steps, 1), np.int64)
        err = self.lib.or3_run_steps(C.byref(self.L), self.n, _ptr(self.is_alive, C.c_int32),
                                     _ptr(self.cell_index, C.c_int32), *self._tables(), C.byref(self.ns),
                                     C.byref(phys), n_steps, int(serial), threads, _ptr(pairs, C.c_int64))
        return err, pairs[:n_steps]

    def term_magnitude(self, what, phys: Physics3D):
        """Sum of |term| over each particle's neighbours: what 0 the continuity terms (n,), 1 the momentum
        terms (n, 3)."""
        out = np.zeros((self.n, 3) if what else self.n, np.float64)
        self.lib.or3_pair_term_magnitude(C.byref(self.L), self.n, _ptr(self.is_alive, C.c_int32), *self._tables(),
                                         C.byref(phys), int(what), _ptr(out, C.c_double))
        return out

    def count(self):
        return self.IP[: self.n, self.L.nCount].copy()

    def neighbours(self, i):
        c = self.IP[i, self.L.nCount]
        return self.IP[i, self.L.nIndex: self.L.nIndex + c].copy()

    def pair_csr(self):
        """(offsets, 0-based neighbour ids in list order) over particles 0..n-1 (dead rows empty)."""
        cnt = np.where(self.is_alive[: self.n] == 1, self.IP[: self.n, self.L.nCount], 0).astype(np.int64)
        off = np.zeros(self.n + 1, np.int64)
        np.cumsum(cnt, out=off[1:])
        M = self.L.max_neighbour_number
        lists = self.IP[: self.n, self.L.nIndex: self.L.nIndex + M] - 1
        mask = np.arange(M)[None, :] < cnt[:, None]
        return off, lists[mask].astype(np.int32)


def build() -> str:
    """Compile the 2D and 3D restatement (one library)."""
    lib_path = O.LIB_PATH
    srcs = [os.path.join(HERE, f) for f in ("eppx_oracle.c", "eppx_oracle3d.c", "eppx_oracle.h")]
    if not os.path.exists(lib_path) or os.path.getmtime(lib_path) < max(os.path.getmtime(s) for s in srcs):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return lib_path
