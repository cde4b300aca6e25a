"""SPH: kernels and the driver's composite actions as libeppx action IDs.

The reference's actions are Julia closures inlined into its kernels
(example/collapse_dry/same/main.jl:74-144 builds them from src/SPH/Library).
A closure cannot cross a C ABI, so each composite the dam-break driver uses is
an action object: an action ID plus an eppx_physics parameter block.  Two
generic field actions cover what the reference's own Action tests do
(scale a column; add a constant per neighbour).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib


class CubicSpline:
    """SPH.Kernel.CubicSpline{IT, FT, 2} (src/SPH/Kernel/CubicSpline.jl)."""
    id = _lib.KERNEL_CUBIC
    sigma = np.float32(10.0 / 7.0 / math.pi)


class WendlandC2:
    """SPH.Kernel.WendlandC2{IT, FT, 2} (src/SPH/Kernel/WendlandC2.jl)."""
    id = _lib.KERNEL_WENDLAND_C2
    sigma = np.float32(7.0 / 4.0 / math.pi)


@dataclass
class Physics:
    """Parameter block of the driver's actions (eppx_physics)."""
    fluid_tag: int = 1
    wall_tag: int = 2
    kernel: int = _lib.KERNEL_WENDLAND_C2
    eos: int = _lib.EOS_TAIT
    visc: int = _lib.VISC_ARTIFICIAL
    dt: float = 0.0
    rho0: float = 1000.0
    p0: float = 0.0
    c0: float = 0.0
    tait_b: float = 0.0
    gx: float = 0.0
    gy: float = -9.81
    mu: float = 1e-3
    alpha: float = 0.1
    beta: float = 0.1
    av_c: float = 0.0
    coefficient: float = 0.0
    gz: float = 0.0            # z component of gravity (3D systems; the 3D sGravity!, State.jl:49-63)

    def c(self) -> _lib.Physics:
        p = _lib.Physics()
        for k in _lib.PHYSICS_INT:
            setattr(p, k, int(getattr(self, k)))
        for k in _lib.PHYSICS_FLOAT:
            setattr(p, k, float(np.float32(getattr(self, k))))
        return p

    def c3(self) -> _lib.Physics3D:
        """The eppx_physics3d block of a 3D system."""
        return _lib.Physics3D(self.c(), float(np.float32(self.gz)))

    def as_dict(self):
        return {k: getattr(self, k) for k in _lib.PHYSICS_INT + _lib.PHYSICS_FLOAT + ("gz",)}


def dam_break_physics(dx: float, h_factor: float, kernel=WendlandC2, g: float = 9.81, column_height: float = 2.0,
                      rho0: float = 1000.0, gamma: int = 7, alpha: float = 0.1, beta: float = 0.1) -> Physics:
    """The benchmark dam break, weakly compressible: Tait EOS (rho0, gamma = 7),
    c0 = 10 sqrt(2 g H), B = rho0 c0^2 / gamma, artificial viscosity
    (alpha, beta, c = c0), dt = 0.1 h / c0; all in Float64, then Float32."""
    c0 = 10.0 * math.sqrt(2.0 * g * column_height)
    h = h_factor * dx
    return Physics(kernel=kernel.id, eos=_lib.EOS_TAIT, visc=_lib.VISC_ARTIFICIAL, dt=float(np.float32(0.1 * h / c0)),
                   rho0=rho0, p0=0.0, c0=float(np.float32(c0)), tait_b=float(np.float32(rho0 * c0 * c0 / gamma)),
                   gx=0.0, gy=-g, alpha=alpha, beta=beta, av_c=float(np.float32(c0)))


def dam_break_physics_3d(dx: float, h_factor: float = 2.0, kernel=WendlandC2, g: float = 9.81,
                         column_height: float = 1.0, rho0: float = 1000.0, gamma: int = 7, alpha: float = 0.1,
                         beta: float = 0.1) -> Physics:
    """The 3D dam break: the constants of dam_break_physics for a column of height column_height, with
    gravity along -z."""
    p = dam_break_physics(dx, h_factor, kernel, g, column_height, rho0, gamma, alpha, beta)
    p.gy, p.gz = 0.0, -g
    return p


def example_physics(dx: float, h_factor: float = 1.5, kernel=CubicSpline, c0: float = 120.0, g: float = 9.8,
                    mu0: float = 1e-3, rho0: float = 1000.0, p0: float = 0.0) -> Physics:
    """The reference example's own constants (example/collapse_dry/same/main.jl:57-70, casegen.jl:71-79):
    linear EOS p = c0^2 (rho - rho0) + p0 with c0 = 120, classic viscosity with mu0 = 1e-3
    (main.jl:110,114 -> Viscosity.jl:10-26), g = 9.8 and dt = parallel(0.1 h0 / c0) * 2 -- h0 and c0
    are Float32 parameters, the product and quotient run in Float64, the result is rounded to Float32
    and then doubled."""
    f32 = np.float32
    h0, c = f32(h_factor * dx), f32(c0)
    dt = f32(f32(0.1 * np.float64(h0) / np.float64(c)) * f32(2))
    return Physics(kernel=kernel.id, eos=_lib.EOS_LINEAR, visc=_lib.VISC_CLASSIC, dt=float(dt), rho0=rho0, p0=p0,
                   c0=float(c), tait_b=0.0, gx=0.0, gy=float(f32(-g)), mu=float(f32(mu0)), alpha=0.0, beta=0.0,
                   av_c=0.0)


# ------------------------------------------------------------------ actions


@dataclass
class Action:
    kind: str                 # "search" | "self" | "inter"
    id: int
    physics: Physics | None = None
    op: tuple | None = None   # (column name, component, value) of a field action
    name: str = ""

    def params(self, ps):
        if self.op is None:
            phys = self.physics or Physics()
            blk = phys.c3() if getattr(ps, "dim", 2) == 3 else phys.c()
            return blk, C.sizeof(blk)
        if getattr(ps, "dim", 2) == 3:
            raise NotImplementedError("field actions are not part of the 3D path")
        col, comp, value = self.op
        blk = _lib.FieldOp(field_of(ps, col), int(comp), float(np.float32(value)))
        return blk, C.sizeof(blk)


def defaultSelfaction():
    return Action("search", _lib.SEARCH_NONE, name="defaultSelfaction!")


def sComputeKernelAndiContinuity(physics: Physics) -> Action:
    """main.jl:74-90: kernel value/gradient and continuity, fused after each neighbour list."""
    return Action("search", _lib.SEARCH_KERNEL_CONTINUITY, physics, name="sComputeKernelAndiContinuity!")


def sContinuity(physics: Physics) -> Action:
    """main.jl:92-105: rho += drho dt, V = m / rho, p = eos(rho)."""
    return Action("self", _lib.SELF_CONTINUITY_EOS, physics, name="sContinuity!")


def iMomentum(physics: Physics) -> Action:
    """main.jl:107-118: pressure + viscosity of fluid particles."""
    return Action("inter", _lib.INTER_MOMENTUM, physics, name="iMomentum!")


def sMomentum(physics: Physics) -> Action:
    """main.jl:120-126: gravity then accelerate-and-move of fluid particles."""
    return Action("self", _lib.SELF_GRAVITY_MOVE, physics, name="sMomentum!")


def iFilter(physics: Physics) -> Action:
    """main.jl:128-134: Shepard sums of fluid neighbours."""
    return Action("inter", _lib.INTER_FILTER, physics, name="iFilter")


def sFilter(physics: Physics) -> Action:
    """main.jl:136-144: Shepard density, volume, pressure."""
    return Action("self", _lib.SELF_FILTER, physics, name="sFilter")


def iExtrapolatePressure(physics: Physics) -> Action:
    """SPH.Library.iExtrapolatePressure! (Pressure.jl:81-101) for wall i and
    fluid j, with w = the pair's kernel value, p0 = physics.p0 and
    g = (physics.gx, physics.gy): SumWeight += w V_j, SumWeightedPressure +=
    w V_j (max(p_j, p0) + max(rho_j g . r_ij, p0))."""
    return Action("inter", _lib.INTER_EXTRAPOLATE_PRESSURE, physics, name="iExtrapolatePressure!")


def sExtrapolatePressure(physics: Physics) -> Action:
    """SPH.Library.sExtrapolatePressure! (Pressure.jl:125-143) on wall
    particles: p = SumWeightedPressure / SumWeight and both cleared when
    SumWeight > 0, else p = p0."""
    return Action("self", _lib.SELF_EXTRAPOLATE_PRESSURE, physics, name="sExtrapolatePressure!")


def ScaleField(column: str, component: int, value: float) -> Action:
    """field(I) *= value for every alive particle."""
    return Action("self", _lib.SELF_SCALE_FIELD, op=(column, component, value), name=f"scale {column}")


def AddPerNeighbour(column: str, component: int, value: float) -> Action:
    """field(I) += value once per neighbour of I."""
    return Action("inter", _lib.INTER_ADD_PER_NEIGHBOUR, op=(column, component, value), name=f"add {column}")


_FIELD_BY_BASIC = {"PositionVec": _lib.FIELD_POSITION, "VelocityVec": _lib.FIELD_VELOCITY,
                   "dVelocityVec": _lib.FIELD_DVELOCITY, "Mass": _lib.FIELD_MASS, "Volume": _lib.FIELD_VOLUME,
                   "Density": _lib.FIELD_DENSITY, "dDensity": _lib.FIELD_DDENSITY, "Pressure": _lib.FIELD_PRESSURE}


def field_of(ps, column: str) -> int:
    """Field enum of a named column (honouring basic_index_map_dict renames)."""
    if column in _FIELD_BY_BASIC:
        return _FIELD_BY_BASIC[column]
    for basic, renamed in ps.basic_index_map_dict.items():
        if renamed == column and basic in _FIELD_BY_BASIC:
            return _FIELD_BY_BASIC[basic]
    raise NotImplementedError(f"column {column} is not a field libeppx actions address")
