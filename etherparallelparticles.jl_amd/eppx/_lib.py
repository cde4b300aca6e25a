"""ctypes binding of libeppx.so (include/eppx.h).

The library is the product: nothing here computes; every call goes through
the C ABI.  Loading fails loudly when libeppx.so has not been built.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

from . import _build

EPPX_OK, EPPX_E_ARG, EPPX_E_HIP, EPPX_E_CAPACITY, EPPX_E_NONFINITE, EPPX_E_STATE, EPPX_E_NOMEM = range(7)
STATUS_NAMES = {0: "EPPX_OK", 1: "EPPX_E_ARG", 2: "EPPX_E_HIP", 3: "EPPX_E_CAPACITY", 4: "EPPX_E_NONFINITE",
                5: "EPPX_E_STATE", 6: "EPPX_E_NOMEM"}

CRIT_DEFAULT, CRIT_SYMMETRY = 0, 1
KERNEL_CUBIC, KERNEL_WENDLAND_C2 = 0, 1
EOS_LINEAR, EOS_TAIT = 0, 1
VISC_CLASSIC, VISC_ARTIFICIAL = 0, 1
SEARCH_NONE, SEARCH_KERNEL_CONTINUITY = 0, 1
SELF_CONTINUITY_EOS, SELF_GRAVITY_MOVE, SELF_FILTER, SELF_EXTRAPOLATE_PRESSURE, SELF_SCALE_FIELD = 1, 2, 3, 4, 10
INTER_MOMENTUM, INTER_FILTER, INTER_EXTRAPOLATE_PRESSURE, INTER_ADD_PER_NEIGHBOUR = 1, 2, 3, 10
(FIELD_POSITION, FIELD_VELOCITY, FIELD_DVELOCITY, FIELD_MASS, FIELD_VOLUME, FIELD_DENSITY, FIELD_DDENSITY,
 FIELD_PRESSURE) = range(8)
MATH_EXACT, MATH_FAST = 0, 1


class EppxError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Domain2D(C.Structure):
    """eppx_domain2d: field order of struct Domain2D (Domain2D.jl:10-26)."""
    _fields_ = [(n, C.c_float) for n in ("gap", "gap_square")] + \
               [(n, C.c_int32) for n in ("n_x", "n_y", "n")] + \
               [(n, C.c_float) for n in ("first_x", "last_x", "first_y", "last_y", "span_x", "span_y",
                                         "gap_x", "gap_y", "gap_x_inv", "gap_y_inv")]


LAYOUT_FIELDS = ("n_int", "n_float", "max_neighbour_number", "Tag", "IsMovable", "nCount", "nIndex",
                 "PositionVec", "VelocityVec", "dVelocityVec", "Mass", "Volume", "Density", "dDensity",
                 "Pressure", "H", "Gap", "SumWeight", "SumWeightedDensity", "nRVec", "nR", "nW", "nDW", "nHInv",
                 "SumWeightedPressure")


class Layout(C.Structure):
    _fields_ = [(n, C.c_int32) for n in LAYOUT_FIELDS]


PHYSICS_INT = ("fluid_tag", "wall_tag", "kernel", "eos", "visc")
PHYSICS_FLOAT = ("dt", "rho0", "p0", "c0", "tait_b", "gx", "gy", "mu", "alpha", "beta", "av_c", "coefficient")


class Physics(C.Structure):
    _fields_ = [(n, C.c_int32) for n in PHYSICS_INT] + [(n, C.c_float) for n in PHYSICS_FLOAT]


class Domain3D(C.Structure):
    """eppx_domain3d: the Domain2D fields with a z axis."""
    _fields_ = [(n, C.c_float) for n in ("gap", "gap_square")] + \
               [(n, C.c_int32) for n in ("n_x", "n_y", "n_z", "n")] + \
               [(n, C.c_float) for n in ("first_x", "last_x", "first_y", "last_y", "first_z", "last_z", "span_x",
                                         "span_y", "span_z", "gap_x", "gap_y", "gap_z", "gap_x_inv", "gap_y_inv",
                                         "gap_z_inv")]


class Physics3D(C.Structure):
    """eppx_physics3d: eppx_physics and the z component of gravity."""
    _fields_ = [("base", Physics), ("gz", C.c_float)]


class FieldOp(C.Structure):
    _fields_ = [("field", C.c_int32), ("component", C.c_int32), ("value", C.c_float)]


_lib = None


def library_path() -> str:
    return _build.LIB


def lib() -> C.CDLL:
    """Load libeppx.so; raises when it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"libeppx.so not built at {path}; run __graft_entry__.build()")
        L = C.CDLL(path)
        P, i32, i64 = C.POINTER, C.c_int32, C.c_int64
        vp = C.c_void_p
        sig = {
            "eppx_abi_version": ([], i32),
            "eppx_ctx_create": ([i32, P(vp)], i32),
            "eppx_ctx_destroy": ([vp], i32),
            "eppx_last_error": ([vp], C.c_char_p),
            "eppx_sync": ([vp], i32),
            "eppx_set_math_mode": ([vp, i32], i32),
            "eppx_domain2d_init": ([C.c_double] * 5 + [P(Domain2D)], i32),
            "eppx_domain2d_locate": ([P(Domain2D), C.c_float, C.c_float, P(i32), P(i32), P(i32), P(i32)], i32),
            "eppx_ps_create": ([vp, i32, P(Layout), P(vp)], i32),
            "eppx_ps_destroy": ([vp], i32),
            "eppx_ps_upload": ([vp, i32, P(i32), P(i32), P(i32), P(C.c_float)], i32),
            "eppx_ps_download": ([vp, vp, P(i32), P(i32), P(i32), P(i32), P(C.c_float)], i32),
            "eppx_ps_download_noalloc": ([vp, vp, P(i32), P(i32), P(i32), P(i32), P(C.c_float)], i32),
            "eppx_ps_alive_count": ([vp, P(i32)], i32),
            "eppx_ns_create": ([vp, P(Domain2D), P(i32), i32, i32, P(vp)], i32),
            "eppx_ns_destroy": ([vp], i32),
            "eppx_ns_cell_stencil": ([vp, P(i32), P(i32)], i32),
            "eppx_search": ([vp, vp, i32, i32, vp, C.c_size_t], i32),
            "eppx_selfaction": ([vp, i32, vp, C.c_size_t], i32),
            "eppx_interaction": ([vp, vp, i32, vp, C.c_size_t], i32),
            "eppx_step": ([vp, vp, P(Physics), i32, i32, i64], i32),
            "eppx_pair_count": ([vp, vp, P(i64)], i32),
            "eppx_pair_total": ([vp, P(i64), i32], i32),
            "eppx_export_pairs": ([vp, vp, P(i64), P(i32), i64], i32),
            "eppx_check_finite": ([vp], i32),
            "eppx_valu_calibration": ([vp, i32, i32, P(C.c_double)], i32),
            "eppx_pair_fallback_blocks": ([vp, P(i64), i32], i32),
            "eppx_pair_row_fallback": ([vp, P(i64), i32], i32),
            "eppx_set_timing": ([vp, i32], i32),
            "eppx_kernel_time": ([vp, C.c_char_p, P(C.c_double), P(i64), i32], i32),
            "eppx_batch_bytes_per_particle": ([P(i32)], i32),
            "eppx_batch_create": ([vp, P(Layout), i32, i32, P(i32), i32, i32, P(vp)], i32),
            "eppx_batch_destroy": ([vp], i32),
            "eppx_batch_add": ([vp, P(Domain2D), P(Physics), i32, P(i32), P(i32), P(i32), P(C.c_float), i32,
                                P(i32)], i32),
            "eppx_batch_step": ([vp, i32], i32),
            "eppx_batch_steps_left": ([vp, P(i32)], i32),
            "eppx_batch_download": ([vp, i32, P(i32), P(i32), P(i32), P(i32), P(C.c_float)], i32),
            "eppx_batch_remove": ([vp, i32], i32),
            "eppx_batch_info": ([vp, P(i32), P(i32)], i32),
            "eppx_batch_pair_total": ([vp, P(i64), i32], i32),
            "eppx_autotune": ([vp, vp, P(Physics), i32, P(i32), P(C.c_double)], i32),
            "eppx_set_pair_shape": ([vp, P(i32)], i32),
            "eppx_batch_autotune": ([vp, i32, P(i32), P(C.c_double)], i32),
            "eppx_domain3d_init": ([C.c_double] * 7 + [P(Domain3D)], i32),
            "eppx_domain3d_locate": ([P(Domain3D), C.c_float, C.c_float, C.c_float, P(i32), P(i32), P(i32), P(i32),
                                      P(i32)], i32),
            "eppx_ns3d_create": ([vp, P(Domain3D), P(i32), i32, i32, P(vp)], i32),
            "eppx_ns3d_destroy": ([vp], i32),
            "eppx_ns3d_cell_stencil": ([vp, P(i32), P(i32)], i32),
            "eppx_ps3d_create": ([vp, i32, P(Layout), P(vp)], i32),
            "eppx_ps3d_destroy": ([vp], i32),
            "eppx_ps3d_upload": ([vp, i32, P(i32), P(i32), P(i32), P(C.c_float)], i32),
            "eppx_ps3d_download": ([vp, vp, P(i32), P(i32), P(i32), P(i32), P(C.c_float)], i32),
            "eppx_ps3d_alive_count": ([vp, P(i32)], i32),
            "eppx_search3d": ([vp, vp, i32, i32, vp, C.c_size_t], i32),
            "eppx_selfaction3d": ([vp, i32, vp, C.c_size_t], i32),
            "eppx_interaction3d": ([vp, vp, i32, vp, C.c_size_t], i32),
            "eppx_step3d": ([vp, vp, P(Physics3D), i32], i32),
            "eppx_pair_count3d": ([vp, vp, P(i64)], i32),
            "eppx_pair_total3d": ([vp, P(i64), i32], i32),
            "eppx_export_pairs3d": ([vp, vp, P(i64), P(i32), i64], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


# Live handles, closed at interpreter exit in dependency order (systems
# before their contexts) while the HIP runtime is still up.
_live = weakref.WeakSet()
_shut_down = False


def track(obj) -> None:
    _live.add(obj)


def shutting_down() -> bool:
    return _shut_down


@atexit.register
def _close_all() -> None:
    global _shut_down
    objs = list(_live)
    for o in objs:
        if getattr(o, "is_context", False) is False:
            o.close()
    for o in objs:
        o.close()
    _shut_down = True


def check(status: int, ctx=None, what: str = "") -> None:
    if status != EPPX_OK:
        msg = lib().eppx_last_error(ctx).decode() if ctx else ""
        raise EppxError(status, f"{what}: {msg}" if what else msg)


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
