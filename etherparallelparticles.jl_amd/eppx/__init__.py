"""eppx: EtherParallelParticles.jl's SPH time step on MI355X (gfx950) through libeppx.

Host-side mirror of the reference's module layout:

* ``Environment`` -- Parallel (device context, IT / FT conversion)
* ``Class``       -- Domain2D, NamedIndex, (Host)ParticleSystem, NeighbourSystem, mirror / syncto / asyncto
* ``Algorithm``   -- search / selfaction / interaction (+ static_ / dynamic_), step
* ``SPH``         -- kernels and the dam-break driver's composite actions
* ``Geometry``    -- Rectangle lattice generator
* ``cases``       -- dam-break cases and benchmark workloads
* ``DataIO``      -- JSON configuration to handles and back, raw output (.npz)
* ``Batch``       -- many independent cases in one set of device buffers

Every compute call goes through include/eppx.h; there is no CPU fallback.
"""
from . import Algorithm, Batch, Class, DataIO, Environment, Geometry, SPH, cases  # noqa: F401
from ._build import build  # noqa: F401
from ._lib import EppxError, library_path, lib  # noqa: F401

__all__ = ["Algorithm", "Batch", "Class", "DataIO", "Environment", "Geometry", "SPH", "cases", "build", "EppxError", "lib",
           "library_path"]
