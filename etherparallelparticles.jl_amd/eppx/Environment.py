"""Environment: the Parallel{IT, FT, CT, Backend} handle (src/Environment/Parallel.jl:13-137).

Here the backend is always the HIP device behind libeppx; a Parallel owns one
eppx context (device + stream), created on first device use so that host-only
work (layouts, case generation) runs without a GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class Parallel:
    """Parallel{Int32, Float32, HIP, device}: type conversion and the device context."""

    is_context = True

    def __init__(self, IT=np.int32, FT=np.float32, device: int = 0, math: str = "exact"):
        if np.dtype(IT) != np.int32 or np.dtype(FT) != np.float32:
            raise NotImplementedError("libeppx computes in Int32 / Float32 (the reference's IT / FT)")
        self.IT, self.FT, self.device = np.int32, np.float32, int(device)
        self._ctx = None
        self._math = math

    # parallel(x) conversion, Parallel.jl:49-101
    def __call__(self, x):
        if isinstance(x, dict):
            return {k: self(v) if isinstance(v, (int, float, np.integer, np.floating)) else v for k, v in x.items()}
        if isinstance(x, (bool, np.bool_)):
            return x
        if isinstance(x, (int, np.integer)):
            return self.IT(x)
        if isinstance(x, (float, np.floating)):
            return self.FT(x)
        a = np.asarray(x)
        if np.issubdtype(a.dtype, np.integer):
            return a.astype(self.IT)
        return a.astype(self.FT)

    @property
    def ctx(self):
        if self._ctx is None:
            L = _lib.lib()
            h = C.c_void_p()
            st = L.eppx_ctx_create(self.device, C.byref(h))
            if st != _lib.EPPX_OK:
                raise _lib.EppxError(st, f"cannot open HIP device {self.device}")
            self._ctx = h
            _lib.track(self)
            self.set_math(self._math)
        return self._ctx

    def set_math(self, mode: str) -> None:
        """'exact': reference operation order; 'fast': FMA + hardware reciprocals in pair terms."""
        self._math = mode
        if self._ctx is not None:
            m = _lib.MATH_FAST if mode == "fast" else _lib.MATH_EXACT
            _lib.check(_lib.lib().eppx_set_math_mode(self._ctx, m), self._ctx, "set_math")

    def synchronize(self) -> None:
        """Environment.synchronize (Parallel.jl:104-108); raises on sticky capacity flags."""
        if self._ctx is not None:
            _lib.check(_lib.lib().eppx_sync(self._ctx), self._ctx, "synchronize")

    def close(self) -> None:
        """Release the context (freed once its particle / neighbour systems are gone)."""
        if self._ctx is not None and not _lib.shutting_down():
            _lib.lib().eppx_ctx_destroy(self._ctx)
        self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __repr__(self):
        return f"Parallel{{Int32, Float32, HIP, device={self.device}}}"


def get_int_type(p: Parallel):
    return p.IT


def get_float_type(p: Parallel):
    return p.FT


def synchronize(p: Parallel) -> None:
    p.synchronize()
