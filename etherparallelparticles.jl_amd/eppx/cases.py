"""Dam-break cases (example/collapse_dry/same/casegen.jl) and the benchmark workloads.

``dam_break`` restates casegen.jl:17-160 with particle_gap = dx, H = Gap *
h_factor and a cell size / domain margin of 2h (casegen's 3dx is 2h for
h = 1.5dx).  Workloads:

* ``SMALL``  : 1.0 x 2.0 m column in a 4 x 3 m tank, dx = 0.02, cubic spline,
               h = 1.5dx (6,518 particles), 1000 steps;
* ``LARGE``  : same geometry at dx = 0.001, Wendland C2, h = 2dx
               (2,030,018 particles), 200 steps;
* ``mix_cases(0)`` : 64 cases, dx in {0.02, 0.01, 0.005, 0.0025}, width
               U[0.5, 1.5], aspect U[0.5, 2], Wendland C2, h = 2dx, 500 steps;
* ``serving_stream(1)`` : Poisson arrivals, N log-uniform in [1e4, 1e6],
               dx = sqrt(2 / N), steps U{200..2000}.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import Geometry, SPH
from .Class import Domain2D, Domain3D, HostParticleSystem, NamedIndex

FLUID_TAG, WALL_TAG = 1, 2
ACTIVE_PAIR = [(FLUID_TAG, FLUID_TAG), (FLUID_TAG, WALL_TAG), (WALL_TAG, FLUID_TAG)]


def dam_break_tables(M: int):
    """int / float named tuples of casegen.jl:48-70 (2D)."""
    int_nt = {"Tag": 1, "IsMovable": 1, "nCount": 1, "nIndex": M}
    float_nt = {"PositionVec": 2, "VelocityVec": 2, "dVelocityVec": 2, "AccelerationVec": 2, "Mass": 1,
                "Volume": 1, "Density": 1, "dDensity": 1, "Pressure": 1, "Gap": 1, "H": 1, "SumWeight": 1,
                "SumWeightedDensity": 1, "SumWeightedPressure": 1, "nW": M, "nDW": M, "nHInv": M,
                "nRVec": 2 * M, "nR": M}
    return int_nt, float_nt


@dataclass
class DamBreakCase:
    dx: float
    h_factor: float
    kernel: type
    M: int
    domain: Domain2D
    host: HostParticleSystem
    physics: SPH.Physics
    int_nt: dict
    float_nt: dict
    basic_parameters: dict
    steps: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n_particles(self):
        return self.host.n_particles_


def dam_break(dx: float, h_factor: float = 2.0, kernel=SPH.WendlandC2, M: int = 64, fluid_width: float = 1.0,
              fluid_height: float = 2.0, box_width: float = 4.0, box_height: float = 3.0, rho0: float = 1000.0,
              p0: float = 0.0, wall_layers: int = 3, steps: int = 0) -> DamBreakCase:
    """Builds the dam-break host particle system exactly as casegen.jl does."""
    x0 = y0 = 0.0
    wall_width = wall_layers * dx
    domain_gap = 2.0 * h_factor * dx
    domain = Domain2D(domain_gap, x0 - domain_gap, y0 - domain_gap, x0 + box_width + domain_gap,
                      y0 + box_height + domain_gap)
    int_nt, float_nt = dam_break_tables(M)
    basic_parameters = {"rho0": rho0, "p0": p0, "FLUID_TAG": FLUID_TAG, "WALL_TAG": WALL_TAG, "gap0": dx,
                        "h0": h_factor * dx}
    named_index = NamedIndex(int_nt, float_nt)
    rects = [  # merge order of casegen.jl:145: fluid, left, right, bottom
        (Geometry.Rectangle(x0, y0, x0 + fluid_width, y0 + fluid_height), FLUID_TAG, 1),
        (Geometry.Rectangle(x0 - wall_width, y0, x0, y0 + box_height), WALL_TAG, 0),
        (Geometry.Rectangle(x0 + box_width, y0, x0 + box_width + wall_width, y0 + box_height), WALL_TAG, 0),
        (Geometry.Rectangle(x0 - wall_width, y0 - wall_width, x0 + box_width + wall_width, y0), WALL_TAG, 0),
    ]
    parts = []
    for rect, tag, movable in rects:
        n = Geometry.count(dx, rect)
        ps = HostParticleSystem(n, named_index, basic_parameters, n_particles=n)
        pos, vol, gap = Geometry.create(dx, rect)
        ps["PositionVec"] = pos
        ps["Volume"] = vol
        ps["Gap"] = gap
        ps["Tag"] = tag
        ps["IsMovable"] = movable
        ps["Density"] = np.float32(rho0)
        ps["Mass"] = ps["Density"] * ps["Volume"]
        ps["H"] = (ps["Gap"].astype(np.float64) * h_factor).astype(np.float32)
        ps["Pressure"] = np.float32(p0)
        parts.append(ps)
    host = parts[0].merge(*parts[1:])
    host.set_is_alive()
    physics = SPH.dam_break_physics(dx, h_factor, kernel, column_height=2.0, rho0=rho0)
    return DamBreakCase(dx, h_factor, kernel, M, domain, host, physics, int_nt, float_nt, basic_parameters, steps,
                        {"fluid_width": fluid_width, "fluid_height": fluid_height})


def small(M: int = 40) -> DamBreakCase:
    """1 x 2 m column, dx = 0.02, cubic spline h = 1.5dx: 6,518 particles, 1000 steps."""
    return dam_break(0.02, 1.5, SPH.CubicSpline, M=M, steps=1000)


def large(M: int = 64) -> DamBreakCase:
    """1 x 2 m column, dx = 0.001, Wendland C2 h = 2dx: 2,030,018 particles, 200 steps."""
    return dam_break(0.001, 2.0, SPH.WendlandC2, M=M, steps=200)


def dam_break_3d_tables(M: int):
    """The 3D named tuples: casegen.jl:48-70's columns with 3-wide vectors and 3 floats per nRVec pair.
    M = 0 gives the neighbour columns width 0 (the device never stores neighbour lists; the neighbour
    system's max_neighbour_number still bounds the counts), which keeps the host tables of a
    million-particle case small."""
    int_nt = {"Tag": 1, "IsMovable": 1, "nCount": 1, "nIndex": M}
    float_nt = {"PositionVec": 3, "VelocityVec": 3, "dVelocityVec": 3, "Mass": 1, "Volume": 1, "Density": 1,
                "dDensity": 1, "Pressure": 1, "Gap": 1, "H": 1, "nW": M, "nDW": M, "nHInv": M, "nRVec": 3 * M,
                "nR": M}
    return int_nt, float_nt


def dam_break_3d(dx: float, h_factor: float = 2.0, kernel=SPH.WendlandC2, M: int = 0, block=(1.0, 0.5, 1.0),
                 tank=(3.2, 1.0, 1.8), rho0: float = 1000.0, p0: float = 0.0, wall_layers: int = 3,
                 steps: int = 0) -> DamBreakCase:
    """The 3D dam break: a water block (block = (x, y, z) extents) in the corner x = y = z = 0 of a tank
    (tank = extents) with wall_layers layers of fixed wall particles on the floor and the four sides,
    Tait EOS, artificial viscosity, gravity along -z, H = Gap * h_factor, cell size and domain margin 2h.
    The reference has no 3D case generator (Geometry3D.jl:89-94 stops at the counts); this one follows
    casegen.jl's 2D construction: fluid first, then walls, each part a lattice of its own, merged."""
    bx, by, bz = block
    tx, ty, tz = tank
    w = wall_layers * dx
    margin = 2.0 * h_factor * dx
    domain = Domain3D(margin, -margin, -margin, -margin, tx + margin, ty + margin, tz + margin)
    int_nt, float_nt = dam_break_3d_tables(M)
    basic_parameters = {"rho0": rho0, "p0": p0, "FLUID_TAG": FLUID_TAG, "WALL_TAG": WALL_TAG, "gap0": dx,
                        "h0": h_factor * dx}
    named_index = NamedIndex(int_nt, float_nt)
    parts_spec = [
        (Geometry.Cuboid(0.0, 0.0, 0.0, bx, by, bz), FLUID_TAG, 1),
        (Geometry.Cuboid(-w, -w, -w, tx + w, ty + w, 0.0), WALL_TAG, 0),    # floor
        (Geometry.Cuboid(-w, -w, 0.0, 0.0, ty + w, tz), WALL_TAG, 0),       # x = 0 side
        (Geometry.Cuboid(tx, -w, 0.0, tx + w, ty + w, tz), WALL_TAG, 0),    # x = tank side
        (Geometry.Cuboid(0.0, -w, 0.0, tx, 0.0, tz), WALL_TAG, 0),          # y = 0 side
        (Geometry.Cuboid(0.0, ty, 0.0, tx, ty + w, tz), WALL_TAG, 0),       # y = tank side
    ]
    parts = []
    for cub, tag, movable in parts_spec:
        n = Geometry.count3(dx, cub)
        ps = HostParticleSystem(n, named_index, basic_parameters, n_particles=n)
        pos, vol, gap = Geometry.create3(dx, cub)
        ps["PositionVec"] = pos
        ps["Volume"] = vol
        ps["Gap"] = gap
        ps["Tag"] = tag
        ps["IsMovable"] = movable
        ps["Density"] = np.float32(rho0)
        ps["Mass"] = ps["Density"] * ps["Volume"]
        ps["H"] = (ps["Gap"].astype(np.float64) * h_factor).astype(np.float32)
        ps["Pressure"] = np.float32(p0)
        parts.append(ps)
    host = parts[0].merge(*parts[1:])
    host.set_is_alive()
    physics = SPH.dam_break_physics_3d(dx, h_factor, kernel, column_height=bz, rho0=rho0)
    return DamBreakCase(dx, h_factor, kernel, M, domain, host, physics, int_nt, float_nt, basic_parameters, steps,
                        {"block": tuple(block), "tank": tuple(tank), "dimension": 3})


def large_3d(M: int = 0) -> DamBreakCase:
    """The 3D dam break at dx = 0.0075: a 1.0 x 0.5 x 1.0 m block (1,185,163 fluid particles) in a
    3.2 x 1.0 x 1.8 m tank with 3 wall layers, Wendland C2 with h = 2dx, 100 steps."""
    return dam_break_3d(0.0075, 2.0, SPH.WendlandC2, M=M, steps=100)


def mix_specs(seed: int = 0, n_cases: int = 64, steps: int = 500):
    """(dx, width, height, steps) of the uneven 64-case mix: for each case in order
    dx ~ {0.02, 0.01, 0.005, 0.0025}, width ~ U[0.5, 1.5], aspect ~ U[0.5, 2]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_cases):
        dx = float(rng.choice([0.02, 0.01, 0.005, 0.0025]))
        width = float(rng.uniform(0.5, 1.5))
        aspect = float(rng.uniform(0.5, 2.0))
        out.append((dx, width, aspect * width, steps))
    return out


def mix_cases(seed: int = 0, n_cases: int = 64, steps: int = 500, M: int = 64):
    return [dam_break(dx, 2.0, SPH.WendlandC2, M=M, fluid_width=w, fluid_height=h, steps=s)
            for dx, w, h, s in mix_specs(seed, n_cases, steps)]


def serving_stream(seed: int = 1, n_requests: int = 32, rate_hz: float = 4.0, n_min: float = 1e4,
                   n_max: float = 1e6, steps_min: int = 200, steps_max: int = 2000):
    """(arrival_s, n_target, dx, steps) of the Poisson serving stream: exponential
    inter-arrivals at rate_hz, N log-uniform in [n_min, n_max], dx = sqrt(2 / N)
    (1 x 2 m column), steps uniform in {steps_min..steps_max}."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for _ in range(n_requests):
        t += float(rng.exponential(1.0 / rate_hz))
        n = float(math.exp(rng.uniform(math.log(n_min), math.log(n_max))))
        dx = math.sqrt(2.0 / n)
        steps = int(rng.integers(steps_min, steps_max + 1))
        out.append((t, n, dx, steps))
    return out


def serving_requests(seed: int = 1, n_requests: int = 32, rate_hz: float = 4.0, n_min: float = 1e4,
                     n_max: float = 1e6, steps_min: int = 200, steps_max: int = 2000, M: int = 1):
    """The serving stream as Serve.Request objects: each a 1 x 2 m column dam
    break at its dx (Wendland C2, h = 2dx, dam_break_physics)."""
    from .Serve import Request
    out = []
    for t, _, dx, steps in serving_stream(seed, n_requests, rate_hz, n_min, n_max, steps_min, steps_max):
        c = dam_break(dx, 2.0, SPH.WendlandC2, M=M, fluid_width=1.0, fluid_height=2.0, steps=steps)
        out.append(Request(t, c.domain, c.physics, c.host, steps, name=f"dx={dx:.5f}"))
    return out
