"""Class: Domain2D, NamedIndex, particle systems and the neighbour system.

Mirrors src/Class of EtherParallelParticles.jl: the host particle system keeps
the reference's column-major (capacity, columns) Int32 / Float32 tables; the
device particle system is an eppx_ps handle whose state lives in HBM as
cell-sorted structure-of-arrays buffers.  Data moves between the two only
through asyncto / syncto, exactly where the reference copies host <-> device.
"""
from __future__ import annotations

import ctypes as C
from collections import OrderedDict

import numpy as np

from . import _lib
from .Environment import Parallel

# ------------------------------------------------------------------ domain


class Domain2D:
    """Domain2D{Int32, Float32}(gap, first_x, first_y, last_x, last_y), Domain2D.jl:28-67.

    The Float32 fields are computed by libeppx's host code (eppx_domain2d_init)
    so the device and the host use the same bits.
    """

    def __init__(self, gap, first_x, first_y, last_x, last_y):
        self.args = (float(gap), float(first_x), float(first_y), float(last_x), float(last_y))
        self.c = _lib.Domain2D()
        _lib.check(_lib.lib().eppx_domain2d_init(*self.args, C.byref(self.c)), None, "Domain2D")

    def __getattr__(self, name):
        if name.endswith("_") and name[:-1] in dict(_lib.Domain2D._fields_):
            return getattr(self.c, name[:-1])
        raise AttributeError(name)

    def dimension(self):
        return 2

    def get_gap(self): return np.float32(self.c.gap)
    def get_gap_square(self): return np.float32(self.c.gap_square)
    def get_n_x(self): return self.c.n_x
    def get_n_y(self): return self.c.n_y
    def get_n(self): return self.c.n
    def get_first_x(self): return np.float32(self.c.first_x)
    def get_first_y(self): return np.float32(self.c.first_y)
    def get_last_x(self): return np.float32(self.c.last_x)
    def get_last_y(self): return np.float32(self.c.last_y)
    def get_span_x(self): return np.float32(self.c.span_x)
    def get_span_y(self): return np.float32(self.c.span_y)
    def get_gap_x(self): return np.float32(self.c.gap_x)
    def get_gap_y(self): return np.float32(self.c.gap_y)
    def get_gap_x_inv(self): return np.float32(self.c.gap_x_inv)
    def get_gap_y_inv(self): return np.float32(self.c.gap_y_inv)

    def _locate(self, x, y):
        ins, i, j, lin = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().eppx_domain2d_locate(C.byref(self.c), float(np.float32(x)), float(np.float32(y)),
                                                  C.byref(ins), C.byref(i), C.byref(j), C.byref(lin)))
        return bool(ins.value), i.value, j.value, lin.value

    # Domain2D.jl:95-137
    def indexCartesianToLinear(self, i, j): return int(i) + self.c.n_x * (int(j) - 1)
    def indexLinearToCartesian(self, index): return (int(index) - 1) % self.c.n_x + 1, -(-int(index) // self.c.n_x)
    def inside(self, x, y): return self._locate(x, y)[0]
    def indexCartesianFromPosition(self, x, y): return self._locate(x, y)[1:3]
    def indexLinearFromPosition(self, x, y): return self._locate(x, y)[3]

    def __repr__(self):
        return (f"Domain2D{{Int32, Float32}}(gap={self.c.gap}, n_x={self.c.n_x}, n_y={self.c.n_y}, "
                f"first=({self.c.first_x}, {self.c.first_y}), last=({self.c.last_x}, {self.c.last_y}))")


class Domain3D:
    """Domain3D{Int32, Float32}(gap, first_x, first_y, first_z, last_x, last_y, last_z): the reference's
    Domain2D (Domain2D.jl:28-67) with a z axis (its 3D domain is a stub, AbstractDomain.jl:119).  Cell
    (i, j, k) is linear i + n_x (j - 1) + n_x n_y (k - 1), 1-based.  Fields computed by eppx_domain3d_init."""

    def __init__(self, gap, first_x, first_y, first_z, last_x, last_y, last_z):
        self.args = tuple(float(v) for v in (gap, first_x, first_y, first_z, last_x, last_y, last_z))
        self.c = _lib.Domain3D()
        _lib.check(_lib.lib().eppx_domain3d_init(*self.args, C.byref(self.c)), None, "Domain3D")

    def __getattr__(self, name):
        if name.endswith("_") and name[:-1] in dict(_lib.Domain3D._fields_):
            return getattr(self.c, name[:-1])
        raise AttributeError(name)

    def dimension(self):
        return 3

    def get_gap(self): return np.float32(self.c.gap)
    def get_n_x(self): return self.c.n_x
    def get_n_y(self): return self.c.n_y
    def get_n_z(self): return self.c.n_z
    def get_n(self): return self.c.n

    def _locate(self, x, y, z):
        ins, i, j, k, lin = (C.c_int32() for _ in range(5))
        _lib.check(_lib.lib().eppx_domain3d_locate(C.byref(self.c), float(np.float32(x)), float(np.float32(y)),
                                                  float(np.float32(z)), C.byref(ins), C.byref(i), C.byref(j),
                                                  C.byref(k), C.byref(lin)))
        return bool(ins.value), i.value, j.value, k.value, lin.value

    def indexCartesianToLinear(self, i, j, k):
        return int(i) + self.c.n_x * ((int(j) - 1) + self.c.n_y * (int(k) - 1))

    def indexLinearToCartesian(self, index):
        c = int(index) - 1
        return c % self.c.n_x + 1, (c // self.c.n_x) % self.c.n_y + 1, c // (self.c.n_x * self.c.n_y) + 1

    def inside(self, x, y, z): return self._locate(x, y, z)[0]
    def indexCartesianFromPosition(self, x, y, z): return self._locate(x, y, z)[1:4]
    def indexLinearFromPosition(self, x, y, z): return self._locate(x, y, z)[4]

    def __repr__(self):
        c = self.c
        return (f"Domain3D{{Int32, Float32}}(gap={c.gap}, n=({c.n_x}, {c.n_y}, {c.n_z}), "
                f"first=({c.first_x}, {c.first_y}, {c.first_z}), last=({c.last_x}, {c.last_y}, {c.last_z}))")


# -------------------------------------------------------------- named index


class NamedIndexTable:
    """NamedIndexTable{IT}(capacity_named_tuple), NamedIndex.jl:50-91."""

    def __init__(self, capacity: dict):
        assert len(capacity) > 0
        self.capacity = OrderedDict((k, int(v)) for k, v in capacity.items())
        self.index = OrderedDict()
        start = 1
        for k, v in self.capacity.items():
            self.index[k] = start
            start += v
        self.n_capacity = start - 1
        head = []
        for k, v in self.capacity.items():
            head += [k] if v == 1 else [k if j == 1 else f"{k}_{j}" for j in range(1, v + 1)]
        self.name_string_table_head = head


class NamedIndex:
    """NamedIndex{IT}(int_named_tuple, float_named_tuple), NamedIndex.jl:93-110."""

    def __init__(self, int_named_tuple: dict, float_named_tuple: dict):
        self.int_table = NamedIndexTable(int_named_tuple)
        self.float_table = NamedIndexTable(float_named_tuple)
        self.combined = OrderedDict(list(self.int_table.index.items()) + list(self.float_table.index.items()))

    def get_n_int_capacity(self): return self.int_table.n_capacity
    def get_n_float_capacity(self): return self.float_table.n_capacity
    def get_int_symbol_head(self): return list(self.int_table.name_string_table_head)
    def get_float_symbol_head(self): return list(self.float_table.name_string_table_head)
    def get_index_named_tuple(self): return dict(self.combined)


kBasicIndexMapDict = {"IsMovable": "IsMovable", "PositionVec": "PositionVec", "Tag": "Tag", "nCount": "nCount",
                      "nIndex": "nIndex", "nRVec": "nRVec", "nR": "nR"}


def mapBasicIndex(named_index: NamedIndex, basic_index_map_dict: dict | None = None) -> dict:
    """mapBasicIndex, BasicIndex.jl:20-44: 1-based column of each basic field after renaming."""
    m = dict(kBasicIndexMapDict)
    for k in kBasicIndexMapDict:
        if basic_index_map_dict and k in basic_index_map_dict:
            m[k] = basic_index_map_dict[k]
    out = {}
    for k, name in m.items():
        if name not in named_index.combined:
            raise AssertionError(f"basic field {k} -> {name} is not a declared property")
        out[k] = named_index.combined[name]
    return out


def _layout(named_index: NamedIndex, basic_index: dict, max_neighbour_number: int | None) -> _lib.Layout:
    """0-based eppx_layout of a particle system: basic fields through the basic
    index, the physics fields by their reference names."""
    L = _lib.Layout()
    L.n_int = named_index.get_n_int_capacity()
    L.n_float = named_index.get_n_float_capacity()
    ints, floats = named_index.int_table.index, named_index.float_table.index
    for f in _lib.LAYOUT_FIELDS[3:]:
        if f in basic_index:
            col = basic_index[f] - 1
        elif f in ints:
            col = ints[f] - 1
        elif f in floats:
            col = floats[f] - 1
        else:
            col = -1
        setattr(L, f, col)
    if max_neighbour_number is None:
        nidx = [k for k, v in ints.items() if v == basic_index.get("nIndex")]
        max_neighbour_number = named_index.int_table.capacity[nidx[0]] if nidx else 0
    L.max_neighbour_number = int(max_neighbour_number)
    return L


# --------------------------------------------------------- particle systems


class _ParticleSystemBase:
    def _init_common(self, named_index, basic_index_map_dict, basic_parameters, n_capacity, n_particles):
        self.named_index_ = named_index
        self.basic_index_ = mapBasicIndex(named_index, basic_index_map_dict)
        self.basic_index_map_dict = dict(basic_index_map_dict or {})
        self.basic_parameters_ = dict(basic_parameters)
        self.parameters_ = {**self.basic_parameters_, **named_index.combined}
        self.n_capacity = int(n_capacity)
        self.n_particles_ = int(n_particles)
        self.layout = _layout(named_index, self.basic_index_, None)

    def get_n_particles(self): return self.n_particles_
    def get_n_capacity(self): return self.n_capacity
    def get_n_int_capacity(self): return self.named_index_.get_n_int_capacity()
    def get_n_float_capacity(self): return self.named_index_.get_n_float_capacity()
    def get_basic_index(self): return dict(self.basic_index_)
    def get_basic_parameters(self): return dict(self.basic_parameters_)


class HostParticleSystem(_ParticleSystemBase):
    """Host particle system (ParticleSystem with the CPU backend, HostParticleSystem.jl)."""

    def __init__(self, n_capacity, named_index: NamedIndex, basic_parameters: dict, basic_index_map_dict=None,
                 n_particles: int = 0):
        self._init_common(named_index, basic_index_map_dict, basic_parameters, n_capacity, n_particles)
        self.is_alive_ = np.zeros(self.n_capacity, np.int32)
        self.cell_index_ = np.zeros(self.n_capacity, np.int32)
        self.int_properties_ = np.zeros((self.n_capacity, self.get_n_int_capacity()), np.int32, order="F")
        self.float_properties_ = np.zeros((self.n_capacity, self.get_n_float_capacity()), np.float32, order="F")

    # name access: get_int / get_float / setindex!, HostParticleSystem.jl:175-335
    def _where(self, name):
        if name in self.named_index_.int_table.index:
            t = self.named_index_.int_table
            return self.int_properties_, t.index[name] - 1, t.capacity[name]
        if name in self.named_index_.float_table.index:
            t = self.named_index_.float_table
            return self.float_properties_, t.index[name] - 1, t.capacity[name]
        raise KeyError(f"unknown property {name}")

    def __getitem__(self, name):
        a, c, w = self._where(name)
        assert self.n_particles_ > 0, "no particles in the system"
        return a[: self.n_particles_, c].copy() if w == 1 else a[: self.n_particles_, c: c + w].copy()

    def __setitem__(self, name, value):
        a, c, w = self._where(name)
        v = np.asarray(value)
        if v.ndim == 0:
            a[: self.n_particles_, c: c + w] = v.astype(a.dtype)
        else:
            n = v.shape[0]
            assert n <= self.n_capacity, "particles number exceed capacity"
            a[:n, c: c + w] = v.reshape(n, -1).astype(a.dtype) if w > 1 else v.reshape(n, 1).astype(a.dtype)

    def set_n_particles(self, n):
        self.n_particles_ = int(n)

    def set_is_alive(self):
        self.is_alive_[: self.n_particles_] = 1

    def reshape(self, n_capacity):
        assert n_capacity > self.n_capacity
        out = HostParticleSystem(n_capacity, self.named_index_, self.basic_parameters_, self.basic_index_map_dict,
                                 self.n_particles_)
        c = self.n_capacity
        out.is_alive_[:c] = self.is_alive_
        out.cell_index_[:c] = self.cell_index_
        out.int_properties_[:c] = self.int_properties_
        out.float_properties_[:c] = self.float_properties_
        return out

    def merge(self, *others):
        """Base.merge, HostParticleSystem.jl:110-134 (rows of later systems appended)."""
        ps = self
        for o in others:
            n1, n2 = ps.n_particles_, o.n_particles_
            new = ps.reshape(ps.n_capacity + o.n_capacity)
            new.is_alive_[n1: n1 + n2] = o.is_alive_[:n2]
            new.cell_index_[n1: n1 + n2] = o.cell_index_[:n2]
            new.int_properties_[n1: n1 + n2] = o.int_properties_[:n2]
            new.float_properties_[n1: n1 + n2] = o.float_properties_[:n2]
            new.n_particles_ = n1 + n2
            ps = new
        return ps

    def column(self, name, component=0):
        a, c, _ = self._where(name)
        return a[: self.n_particles_, c + component]


class ParticleSystem(_ParticleSystemBase):
    """Device particle system: ParticleSystem(parallel, domain, n_particles, [n_capacity,]
    int_named_tuple, float_named_tuple, basic_parameters; capacityExpand,
    basic_index_map_dict), ParticleSystem.jl:46-115."""

    def __init__(self, parallel: Parallel, domain: Domain2D, n_particles, int_named_tuple, float_named_tuple,
                 basic_parameters, n_capacity=None, capacityExpand=None, basic_index_map_dict=None,
                 max_neighbour_number=None):
        if n_capacity is None:
            n_capacity = capacityExpand(n_particles) if capacityExpand else n_particles
        named_index = NamedIndex(int_named_tuple, float_named_tuple)
        self._init_common(named_index, basic_index_map_dict, parallel(dict(basic_parameters)), n_capacity,
                          n_particles)
        if max_neighbour_number is not None:
            self.layout.max_neighbour_number = int(max_neighbour_number)
        self.parallel = parallel
        self.domain = domain
        self.dim = domain.dimension() if domain is not None else 2
        self.last_ns = None
        h = C.c_void_p()
        create = _lib.lib().eppx_ps3d_create if self.dim == 3 else _lib.lib().eppx_ps_create
        _lib.check(create(parallel.ctx, self.n_capacity, C.byref(self.layout), C.byref(h)), parallel.ctx,
                   "ParticleSystem")
        self.handle = h
        _lib.track(self)
        # ParticleSystem.jl:74-77: the first n_particles rows start alive
        host = mirror(self)
        host.is_alive_[: self.n_particles_] = 1
        syncto(self, host)

    def close(self):
        if getattr(self, "handle", None) is not None and not _lib.shutting_down():
            (_lib.lib().eppx_ps3d_destroy if self.dim == 3 else _lib.lib().eppx_ps_destroy)(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_alive_n_particles(self):
        out = C.c_int32()
        fn = _lib.lib().eppx_ps3d_alive_count if self.dim == 3 else _lib.lib().eppx_ps_alive_count
        _lib.check(fn(self.handle, C.byref(out)), self.parallel.ctx, "alive count")
        return out.value

    def set_n_particles(self, n):
        self.n_particles_ = int(n)


def mirror(ps) -> HostParticleSystem:
    """Class.mirror, ParticleSystem.jl:221-273: a host system of the same shape (data copied)."""
    host = HostParticleSystem(ps.n_capacity, ps.named_index_, ps.basic_parameters_, ps.basic_index_map_dict,
                              ps.n_particles_)
    host.layout = ps.layout
    if isinstance(ps, ParticleSystem) and getattr(ps, "_uploaded", False):
        syncto(host, ps)
    elif isinstance(ps, HostParticleSystem):
        syncto(host, ps)
    return host


def syncto(destination, source) -> None:
    """Class.syncto! / asyncto! / to!, ParticleSystem.jl:275-309."""
    if isinstance(destination, ParticleSystem) and isinstance(source, HostParticleSystem):
        destination.n_particles_ = source.n_particles_
        src = source
        ip = np.asfortranarray(src.int_properties_, dtype=np.int32)
        fp = np.asfortranarray(src.float_properties_, dtype=np.float32)
        alive = np.ascontiguousarray(src.is_alive_, dtype=np.int32)
        cell = np.ascontiguousarray(src.cell_index_, dtype=np.int32)
        if ip.shape != (destination.n_capacity, destination.get_n_int_capacity()) or \
                fp.shape != (destination.n_capacity, destination.get_n_float_capacity()):
            raise ValueError("host and device particle systems differ in shape")
        L = _lib.lib()
        upload = L.eppx_ps3d_upload if destination.dim == 3 else L.eppx_ps_upload
        _lib.check(upload(destination.handle, destination.n_particles_, _lib.ptr(alive, C.c_int32),
                          _lib.ptr(cell, C.c_int32), _lib.ptr(ip, C.c_int32), _lib.ptr(fp, C.c_float)),
                   destination.parallel.ctx, "upload")
        destination._uploaded = True
        destination.last_ns = None
    elif isinstance(destination, HostParticleSystem) and isinstance(source, ParticleSystem):
        n = C.c_int32()
        ns = source.last_ns.handle if source.last_ns is not None else None
        L = _lib.lib()
        download = L.eppx_ps3d_download if source.dim == 3 else L.eppx_ps_download
        st = download(source.handle, ns, C.byref(n), _lib.ptr(destination.is_alive_, C.c_int32),
                                _lib.ptr(destination.cell_index_, C.c_int32),
                                _lib.ptr(destination.int_properties_, C.c_int32),
                                _lib.ptr(destination.float_properties_, C.c_float))
        destination.n_particles_ = n.value
        _lib.check(st, source.parallel.ctx, "download")
    elif isinstance(destination, HostParticleSystem) and isinstance(source, HostParticleSystem):
        destination.n_particles_ = source.n_particles_
        destination.is_alive_[:] = source.is_alive_
        destination.cell_index_[:] = source.cell_index_
        destination.int_properties_[:] = source.int_properties_
        destination.float_properties_[:] = source.float_properties_
    else:
        raise TypeError("syncto needs a host and a device particle system")


asyncto = syncto
to = syncto

# -------------------------------------------------------- neighbour system


class AbstractPeriodicBoundaryPolicy:
    """Periodic boundary policies, PeriodicBoundary.jl:10-26.  A policy is a
    class; a periodic one carries its axes in `axis`, as the reference's type
    parameters {X, Y} / {X, Y, Z} do."""

    axis: tuple = ()


class NonePeriodicBoundaryPolicy(AbstractPeriodicBoundaryPolicy):
    """PeriodicBoundary.jl:12: no periodic axis (the policy of every reference example)."""


class HavePeriodicBoundaryPolicy(AbstractPeriodicBoundaryPolicy):
    """PeriodicBoundary.jl:13."""


_PERIODIC_POLICIES: dict = {}


def _periodic_policy(dim: int, axis) -> type:
    axis = tuple(bool(a) for a in axis)
    if len(axis) != dim:
        raise ValueError(f"a {dim}D periodic policy takes {dim} axis flags, got {len(axis)}")
    key = (dim, axis)
    if key not in _PERIODIC_POLICIES:
        base = PeriodicBoundaryPolicy2D if dim == 2 else PeriodicBoundaryPolicy3D
        name = f"{base.__name__}{{{', '.join('true' if a else 'false' for a in axis)}}}"
        _PERIODIC_POLICIES[key] = type(name, (base,), {"axis": axis, "__module__": __name__})
    return _PERIODIC_POLICIES[key]


class PeriodicBoundaryPolicy2D(HavePeriodicBoundaryPolicy):
    """PeriodicBoundaryPolicy2D{X, Y} (PeriodicBoundary.jl:14): PeriodicBoundaryPolicy2D.of(x, y)."""

    @staticmethod
    def of(x: bool, y: bool) -> type:
        return _periodic_policy(2, (x, y))


class PeriodicBoundaryPolicy3D(HavePeriodicBoundaryPolicy):
    """PeriodicBoundaryPolicy3D{X, Y, Z} (PeriodicBoundary.jl:15): PeriodicBoundaryPolicy3D.of(x, y, z)."""

    @staticmethod
    def of(x: bool, y: bool, z: bool) -> type:
        return _periodic_policy(3, (x, y, z))


# the named policies of PeriodicBoundary.jl:17-26
PeriodicBoundaryPolicy2DAlongX = PeriodicBoundaryPolicy2D.of(True, False)
PeriodicBoundaryPolicy2DAlongY = PeriodicBoundaryPolicy2D.of(False, True)
PeriodicBoundaryPolicy2DAlongXY = PeriodicBoundaryPolicy2D.of(True, True)
PeriodicBoundaryPolicy3DAlongX = PeriodicBoundaryPolicy3D.of(True, False, False)
PeriodicBoundaryPolicy3DAlongY = PeriodicBoundaryPolicy3D.of(False, True, False)
PeriodicBoundaryPolicy3DAlongZ = PeriodicBoundaryPolicy3D.of(False, False, True)
PeriodicBoundaryPolicy3DAlongXY = PeriodicBoundaryPolicy3D.of(True, True, False)
PeriodicBoundaryPolicy3DAlongYZ = PeriodicBoundaryPolicy3D.of(False, True, True)
PeriodicBoundaryPolicy3DAlongZX = PeriodicBoundaryPolicy3D.of(True, False, True)
PeriodicBoundaryPolicy3DAlongXYZ = PeriodicBoundaryPolicy3D.of(True, True, True)


def _check_policy(policy) -> type:
    if not (isinstance(policy, type) and issubclass(policy, AbstractPeriodicBoundaryPolicy)) \
            or policy in (AbstractPeriodicBoundaryPolicy, HavePeriodicBoundaryPolicy, PeriodicBoundaryPolicy2D,
                          PeriodicBoundaryPolicy3D):
        raise TypeError(f"not a periodic boundary policy: {policy!r}")
    return policy


class PeriodicBoundary:
    """PeriodicBoundary(parallel, domain, policy), PeriodicBoundary.jl:28-60.

    `neighbour_cell_relative_position_list_` has the reference's shape: (1, 1, 1)
    for NonePeriodicBoundaryPolicy (:36-46), (n_cells, 3^N, N) for a periodic
    policy (:48-60).  The reference fills it with zeros and leaves filling it as a
    TODO (:57); nothing else writes it, so the shift that
    applyPeriodicBoundaryPolicy adds to (dx, dy) (NeighbourSearch.jl:129-139) is
    always zero and a periodic policy finds the pairs and sums of the
    non-periodic search.  The table is kept here, read-only, so that it stays
    what the reference's is; the device search adds no shift."""

    def __init__(self, parallel, domain, periodic_boundary_policy=NonePeriodicBoundaryPolicy):
        self.policy = _check_policy(periodic_boundary_policy)
        if self.policy is NonePeriodicBoundaryPolicy:
            shape = (1, 1, 1)
        else:
            n = domain.dimension()
            shape = (int(domain.get_n()), 3 ** n, n)
        table = np.zeros(shape, np.float32, order="F")
        table.flags.writeable = False
        self.neighbour_cell_relative_position_list_ = table

    def __repr__(self):
        return (f"PeriodicBoundary{{Float32, {self.policy.__name__}}}(shape of "
                f"`neighbour_cell_relative_position_list`: {self.neighbour_cell_relative_position_list_.shape})")


class ActivePair:
    """ActivePair(parallel, pair_vector), ActivePair.jl:17-33."""

    def __init__(self, pair_vector):
        self.pair_vector_ = [(int(a), int(b)) for a, b in pair_vector]
        maximum_tag = max([1] + [max(a, b) for a, b in self.pair_vector_])
        self.adjacency_matrix_ = np.zeros((maximum_tag, maximum_tag), np.int32, order="F")
        for a, b in self.pair_vector_:
            self.adjacency_matrix_[a - 1, b - 1] = 1


class NeighbourSystem:
    """NeighbourSystem(parallel, domain, active_pair, periodic_boundary_policy;
    max_neighbour_number, n_threads), NeighbourSystem.jl:68-86."""

    def __init__(self, parallel: Parallel, domain: Domain2D, active_pair, periodic_boundary_policy=NonePeriodicBoundaryPolicy,
                 max_neighbour_number: int = 50, n_threads: int = 256):
        self.parallel, self.domain = parallel, domain
        self.dim = domain.dimension()
        self.periodic_boundary_ = PeriodicBoundary(parallel, domain, periodic_boundary_policy)
        self.active_pair_ = ActivePair(active_pair)
        self.max_neighbour_number = int(max_neighbour_number)
        adj = np.asfortranarray(self.active_pair_.adjacency_matrix_)
        h = C.c_void_p()
        create = _lib.lib().eppx_ns3d_create if self.dim == 3 else _lib.lib().eppx_ns_create
        _lib.check(create(parallel.ctx, C.byref(domain.c), _lib.ptr(adj, C.c_int32), adj.shape[0],
                          self.max_neighbour_number, C.byref(h)), parallel.ctx, "NeighbourSystem")
        self.handle = h
        _lib.track(self)

    def close(self):
        if getattr(self, "handle", None) is not None and not _lib.shutting_down():
            (_lib.lib().eppx_ns3d_destroy if self.dim == 3 else _lib.lib().eppx_ns_destroy)(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_n(self):
        return self.domain.get_n()

    def neighbour_cell_index(self):
        """(count, list) of device_initializeNeighbourSystem! (NeighbourSystemBase.jl:28-48)."""
        n = self.domain.get_n()
        count = np.zeros(n, np.int32)
        lst = np.zeros((n, 27 if self.dim == 3 else 9), np.int32, order="F")
        fn = _lib.lib().eppx_ns3d_cell_stencil if self.dim == 3 else _lib.lib().eppx_ns_cell_stencil
        _lib.check(fn(self.handle, _lib.ptr(count, C.c_int32), _lib.ptr(lst, C.c_int32)), self.parallel.ctx,
                   "cell stencil")
        return count, lst


def clean(neighbour_system: NeighbourSystem) -> None:
    """Class.clean!(ns): the device search resets its cell counts itself."""


def get_n_particles(ps):
    return ps.get_n_particles()
