"""Batch: many dam-break cases resident in one set of device buffers.

The reference runs one case per process (example/collapse_dry/same/main.jl);
it has no batch API.  This class binds libeppx's eppx_batch_* entry points:
cases of different sizes, domains and step budgets share one segmented cell
list and advance together, one launch per kernel per step (a recorded
hipGraph), each stopping after exactly its own budget.  Cases are added and
removed between steps, which is what a serving loop needs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .Class import Domain2D, HostParticleSystem
from .Environment import Parallel
from .SPH import Physics


def bytes_per_particle() -> int:
    """Device bytes one particle slot of a batch takes (admission by HBM footprint)."""
    out = C.c_int32()
    _lib.check(_lib.lib().eppx_batch_bytes_per_particle(C.byref(out)), None, "bytes_per_particle")
    return out.value


class Batch:
    """Batch(parallel, template, particle_capacity, case_capacity, active_pair, max_neighbour_number).

    `template` is a HostParticleSystem whose named tables every case shares.
    """

    def __init__(self, parallel: Parallel, template: HostParticleSystem, particle_capacity: int, case_capacity: int,
                 active_pair, max_neighbour_number: int = 64):
        from .Class import ActivePair
        self.parallel = parallel
        self.template = template
        self.layout = template.layout
        adj = np.asfortranarray(ActivePair(active_pair).adjacency_matrix_)
        h = C.c_void_p()
        _lib.check(_lib.lib().eppx_batch_create(parallel.ctx, C.byref(self.layout), int(particle_capacity),
                                                int(case_capacity), _lib.ptr(adj, C.c_int32), adj.shape[0],
                                                int(max_neighbour_number), C.byref(h)), parallel.ctx, "Batch")
        self.handle = h
        self.case_capacity = int(case_capacity)
        self.hosts = {}  # case id -> host system shape (for download)
        _lib.track(self)

    def close(self):
        if getattr(self, "handle", None) is not None and not _lib.shutting_down():
            _lib.lib().eppx_batch_destroy(self.handle)
        self.handle = None

    def add(self, domain: Domain2D, physics: Physics, host: HostParticleSystem, steps: int) -> int:
        """Adds a case (its first search runs now); returns its case id."""
        n = host.n_particles_
        ip = np.asfortranarray(host.int_properties_[:n], dtype=np.int32)
        fp = np.asfortranarray(host.float_properties_[:n], dtype=np.float32)
        alive = np.ascontiguousarray(host.is_alive_[:n], dtype=np.int32)
        cell = np.ascontiguousarray(host.cell_index_[:n], dtype=np.int32)
        p = physics.c()
        cid = C.c_int32()
        _lib.check(_lib.lib().eppx_batch_add(self.handle, C.byref(domain.c), C.byref(p), n, _lib.ptr(alive, C.c_int32),
                                             _lib.ptr(cell, C.c_int32), _lib.ptr(ip, C.c_int32),
                                             _lib.ptr(fp, C.c_float), int(steps), C.byref(cid)),
                   self.parallel.ctx, "Batch.add")
        self.hosts[cid.value] = host
        return cid.value

    def step(self, n_steps: int = 1, sync: bool = True) -> None:
        _lib.check(_lib.lib().eppx_batch_step(self.handle, int(n_steps)), self.parallel.ctx, "Batch.step")
        if sync:
            self.parallel.synchronize()

    def steps_left(self) -> np.ndarray:
        out = np.zeros(self.case_capacity, np.int32)
        _lib.check(_lib.lib().eppx_batch_steps_left(self.handle, _lib.ptr(out, C.c_int32)), self.parallel.ctx,
                   "Batch.steps_left")
        return out

    def download(self, case_id: int) -> HostParticleSystem:
        """The case's state as a host system shaped like the one it was added from."""
        src = self.hosts[case_id]
        n = src.n_particles_
        host = HostParticleSystem(n, src.named_index_, src.basic_parameters_, src.basic_index_map_dict, n)
        host.layout = src.layout
        cnt = C.c_int32()
        _lib.check(_lib.lib().eppx_batch_download(self.handle, int(case_id), C.byref(cnt),
                                                  _lib.ptr(host.is_alive_, C.c_int32),
                                                  _lib.ptr(host.cell_index_, C.c_int32),
                                                  _lib.ptr(host.int_properties_, C.c_int32),
                                                  _lib.ptr(host.float_properties_, C.c_float)),
                   self.parallel.ctx, "Batch.download")
        return host

    def remove(self, case_id: int) -> None:
        _lib.check(_lib.lib().eppx_batch_remove(self.handle, int(case_id)), self.parallel.ctx, "Batch.remove")
        self.hosts.pop(case_id, None)

    def autotune(self, trials: int = 3) -> dict:
        """Pair-kernel shapes for the current cases (see Algorithm.autotune)."""
        from .Algorithm import _shape_dict
        choice = (C.c_int32 * 4)()
        us = (C.c_double * 8)()
        _lib.check(_lib.lib().eppx_batch_autotune(self.handle, int(trials), choice, us), self.parallel.ctx,
                   "Batch.autotune")
        return _shape_dict(choice, us)

    def info(self):
        slots, cases = C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().eppx_batch_info(self.handle, C.byref(slots), C.byref(cases)), self.parallel.ctx, "info")
        return slots.value, cases.value

    def pair_total(self, reset: bool = False) -> int:
        out = C.c_int64()
        _lib.check(_lib.lib().eppx_batch_pair_total(self.handle, C.byref(out), int(reset)), self.parallel.ctx,
                   "Batch.pair_total")
        return out.value
