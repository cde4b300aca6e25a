"""Algorithm: search! / selfaction! / interaction! over libeppx (src/Algorithm).

Same call surface as the reference (NeighbourSearch.jl:323-413,
Action.jl:53-206).  Like the reference, each call waits for the device and
then reports (synchronize after every launch); pass sync=False to leave the
work queued.  static_* and dynamic_* exist for source compatibility: the HIP
kernels are compiled once, so the two differ in nothing.
"""
from __future__ import annotations

import ctypes as C

from . import SPH, _lib
from .Class import Domain2D, NeighbourSystem, ParticleSystem

kDefaultThreadNumber = 256
kDefaultMaxNeighbourNumber = 50


class _Criterion:
    def __init__(self, cid, name):
        self.id, self.name = cid, name

    def __repr__(self):
        return self.name


defaultSearchCriterion = _Criterion(_lib.CRIT_DEFAULT, "defaultSearchCriterion")      # NeighbourSearch.jl:141-151
symmetrySearchCriterion = _Criterion(_lib.CRIT_SYMMETRY, "symmetrySearchCriterion")   # NeighbourSearch.jl:153-164


def defaultSelfaction():
    return SPH.defaultSelfaction()


def _criterion(criterion) -> int:
    if isinstance(criterion, _Criterion):
        return criterion.id
    raise NotImplementedError("only defaultSearchCriterion and symmetrySearchCriterion cross the C ABI")


def _action(action, kind: str) -> SPH.Action:
    if action is None:
        action = SPH.defaultSelfaction()
    if not isinstance(action, SPH.Action):
        raise NotImplementedError("Julia-closure actions cannot cross the C ABI; use the eppx.SPH action menu")
    if action.kind != kind:
        raise ValueError(f"{action.name} is a {action.kind} action, not a {kind} action")
    return action


def search(particle_system: ParticleSystem, domain: Domain2D, neighbour_system: NeighbourSystem, *,
           criterion=defaultSearchCriterion, action=None, n_threads: int = kDefaultThreadNumber, sync: bool = True):
    """search!(ps, domain, ns; criterion, action!, n_threads)."""
    act = _action(action, "search")
    blk, size = act.params(particle_system)
    ctx = particle_system.parallel.ctx
    fn = _lib.lib().eppx_search3d if particle_system.dim == 3 else _lib.lib().eppx_search
    _lib.check(fn(particle_system.handle, neighbour_system.handle, _criterion(criterion), act.id, C.byref(blk), size),
               ctx, "search")
    particle_system.last_ns = neighbour_system
    if sync:
        particle_system.parallel.synchronize()


static_search = search
dynamic_search = search


def selfaction(particle_system: ParticleSystem, action, *, n_threads: int = kDefaultThreadNumber, sync: bool = True):
    """selfaction!(ps, action!; n_threads)."""
    act = _action(action, "self")
    blk, size = act.params(particle_system)
    ctx = particle_system.parallel.ctx
    fn = _lib.lib().eppx_selfaction3d if particle_system.dim == 3 else _lib.lib().eppx_selfaction
    _lib.check(fn(particle_system.handle, act.id, C.byref(blk), size), ctx, "selfaction")
    if sync:
        particle_system.parallel.synchronize()


static_selfaction = selfaction
dynamic_selfaction = selfaction


def interaction(particle_system: ParticleSystem, action, *, n_threads: int = kDefaultThreadNumber,
                sync: bool = True):
    """interaction!(ps, action!; n_threads) over the neighbour lists of the last search."""
    act = _action(action, "inter")
    if particle_system.last_ns is None:
        raise _lib.EppxError(_lib.EPPX_E_STATE, "interaction before any search")
    blk, size = act.params(particle_system)
    ctx = particle_system.parallel.ctx
    fn = _lib.lib().eppx_interaction3d if particle_system.dim == 3 else _lib.lib().eppx_interaction
    _lib.check(fn(particle_system.handle, particle_system.last_ns.handle, act.id, C.byref(blk), size), ctx,
               "interaction")
    if sync:
        particle_system.parallel.synchronize()


static_interaction = interaction
dynamic_interaction = interaction


def step(particle_system: ParticleSystem, neighbour_system: NeighbourSystem, physics: SPH.Physics, n_steps: int = 1,
         filter_interval: int = 0, step_offset: int = 0, sync: bool = True):
    """n_steps of the dam-break loop body (main.jl:166-183) in one call."""
    ctx = particle_system.parallel.ctx
    if particle_system.dim == 3:
        if filter_interval:
            raise NotImplementedError("the Shepard filter is not part of the 3D path")
        p3 = physics.c3()
        _lib.check(_lib.lib().eppx_step3d(particle_system.handle, neighbour_system.handle, C.byref(p3), int(n_steps)),
                   ctx, "step")
        if sync:
            particle_system.parallel.synchronize()
        return
    p = physics.c()
    _lib.check(_lib.lib().eppx_step(particle_system.handle, neighbour_system.handle, C.byref(p), int(n_steps),
                                    int(filter_interval), int(step_offset)), ctx, "step")
    if sync:
        particle_system.parallel.synchronize()


SHAPES = ((256, 0), (512, 0), (256, 1), (512, 1))  # (threads per block, per-neighbour data staged in LDS)


def _shape_dict(choice, us):
    return {"search": {"block": choice[0], "staged": bool(choice[1])},
            "momentum": {"block": choice[2], "staged": bool(choice[3])},
            "search_us": {f"{b}{'s' if st else 'g'}": us[k] for k, (b, st) in enumerate(SHAPES)},
            "momentum_us": {f"{b}{'s' if st else 'g'}": us[4 + k] for k, (b, st) in enumerate(SHAPES)}}


def autotune(particle_system: ParticleSystem, neighbour_system: NeighbourSystem, physics: SPH.Physics,
             trials: int = 3) -> dict:
    """Picks the pair-kernel shapes of the search action and the momentum pass (256 or 512 threads per block;
    per-neighbour data gathered from HBM or staged in LDS) by timing each on the current state, which it leaves
    unchanged; FAST math only.  Returns the choice and the timings in microseconds (keys 256g, 512g, 256s, 512s)."""
    if particle_system.dim == 3:
        raise NotImplementedError("the 3D pair kernels have one shape")
    p = physics.c()
    choice = (C.c_int32 * 4)()
    us = (C.c_double * 8)()
    ctx = particle_system.parallel.ctx
    _lib.check(_lib.lib().eppx_autotune(particle_system.handle, neighbour_system.handle, C.byref(p), int(trials),
                                        choice, us), ctx, "autotune")
    return _shape_dict(choice, us)


def set_pair_shape(particle_system: ParticleSystem, search=(256, 0), momentum=(256, 0)) -> None:
    """Pins the pair-kernel shapes ((block, staged) each); results do not depend on them."""
    choice = (C.c_int32 * 4)(search[0], int(search[1]), momentum[0], int(momentum[1]))
    _lib.check(_lib.lib().eppx_set_pair_shape(particle_system.handle, choice), particle_system.parallel.ctx,
               "set_pair_shape")
