"""Python view of one engine context (include/axbc.h) on one GPU.

Mirrors the reference's device interface (src/bc2d.h:136-190): a context is
initcuda() for one graph, pivot() is one bc_func_mono() round, and the
getters return the FP64 state.  All compute runs in libaxbc_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .graph import Graph


class EngineError(RuntimeError):
    pass


class BudgetError(EngineError):
    """An allocation would take the context past its device memory budget
    (AXBC_EBUDGET); the context stays usable."""


AXBC_EBUDGET = -5


def _check(rc: int) -> None:
    if rc != 0:
        msg = _ffi.hip().axbc_last_error()
        text = f"axbc error {rc}: {msg.decode() if msg else ''}"
        raise BudgetError(text) if rc == AXBC_EBUDGET else EngineError(text)


def device_count() -> int:
    return int(_ffi.hip().axbc_device_count())


class Engine:
    """A resident graph plus the per-vertex Brandes state on one device."""

    def __init__(self, g: Graph, device: int = 0):
        L = _ffi.hip()
        if L.axbc_device_count() <= 0:
            raise EngineError("no HIP device visible: the engine runs only on the GPU")
        self.g = g
        self.N = g.N
        ctx = C.c_void_p()
        nbytes = C.c_size_t(0)
        reach = np.ascontiguousarray(g.reach, dtype=np.uint32)
        self._reach = reach
        _check(L.axbc_ctx_create(device, g.N, g.ned, np.ascontiguousarray(g.col, np.uint32),
                                 np.ascontiguousarray(g.row, np.uint32) if g.ned else
                                 np.zeros(1, np.uint32), reach.ctypes.data, g.heuristic,
                                 C.byref(ctx), C.byref(nbytes)))
        self.ctx = ctx
        self.device_bytes = int(nbytes.value)

    def set_budget(self, nbytes: int) -> None:
        """Device bytes the context may hold; allocations past it raise
        BudgetError (0: no limit)."""
        _check(_ffi.hip().axbc_set_memory_budget(self.ctx, int(nbytes)))

    def bytes_in_use(self) -> int:
        n = C.c_size_t(0)
        _check(_ffi.hip().axbc_device_bytes(self.ctx, C.byref(n)))
        return int(n.value)

    def close(self) -> None:
        if getattr(self, "ctx", None):
            _ffi.hip().axbc_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- one round ----------------------------------------------------------
    def pivot(self, v0: int, p: float, stats: bool = False) -> dict | None:
        st = _ffi.PivotStats()
        _check(_ffi.hip().axbc_pivot(self.ctx, int(v0), float(p), C.byref(st) if stats else None))
        return st.as_dict() if stats else None

    def set_batch_width(self, width: int) -> None:
        _check(_ffi.hip().axbc_set_batch_width(self.ctx, int(width)))

    def set_batch_mode(self, mode: int) -> None:
        """0 automatic, 1 bit-packed sweep, 2 one workgroup per root."""
        _check(_ffi.hip().axbc_set_batch_mode(self.ctx, int(mode)))

    def batch(self, v0s, ps, stop_c: int = 0, stats: bool = False):
        """Multi-source rounds (bit-packed, or one workgroup per root): the
        same result as one pivot() per root in order.  Returns (applied,
        stats or None)."""
        L = _ffi.hip()
        v = np.ascontiguousarray(v0s, dtype=np.uint32)
        p = np.ascontiguousarray(ps, dtype=np.float64)
        st = (_ffi.PivotStats * len(v))()
        ap = C.c_int(0)
        _check(L.axbc_batch(self.ctx, v, p, len(v), int(stop_c), C.byref(ap),
                            C.cast(st, C.c_void_p) if stats else None))
        return ap.value, ([s.as_dict() for s in st] if stats else None)

    def batch_delta(self, k: int) -> np.ndarray:
        out = np.empty(self.N, np.float64)
        _check(_ffi.hip().axbc_batch_delta(self.ctx, int(k), out))
        return out

    def batch_levels(self, k: int) -> tuple[np.ndarray, np.ndarray]:
        """Root k's levels and path counts of the last batch."""
        lvl = np.empty(self.N, np.int32)
        sig = np.empty(self.N, np.uint32)
        _check(_ffi.hip().axbc_batch_levels(self.ctx, int(k), lvl, sig))
        return lvl, sig

    def sync(self) -> None:
        _check(_ffi.hip().axbc_sync(self.ctx))

    def last_stats(self) -> dict:
        st = _ffi.PivotStats()
        _check(_ffi.hip().axbc_last_stats(self.ctx, C.byref(st)))
        return st.as_dict()

    def last_timing(self) -> tuple[float, float, float]:
        f, b, u = C.c_float(), C.c_float(), C.c_float()
        _check(_ffi.hip().axbc_last_timing(self.ctx, C.byref(f), C.byref(b), C.byref(u)))
        return f.value, b.value, u.value

    # -- state --------------------------------------------------------------
    def bc(self) -> np.ndarray:
        out = np.empty(self.N, np.float64)
        _check(_ffi.hip().axbc_get_bc(self.ctx, out))
        return out

    def delta(self) -> np.ndarray:
        out = np.empty(self.N, np.float64)
        _check(_ffi.hip().axbc_get_delta(self.ctx, out))
        return out

    def lvl(self) -> np.ndarray:
        out = np.empty(self.N, np.int32)
        _check(_ffi.hip().axbc_get_lvl(self.ctx, out))
        return out

    def sigma(self) -> np.ndarray:
        out = np.empty(self.N, np.uint32)
        _check(_ffi.hip().axbc_get_sigma(self.ctx, out))
        return out

    def deg(self) -> np.ndarray:
        out = np.empty(self.N, np.uint32)
        _check(_ffi.hip().axbc_get_deg(self.ctx, out))
        return out

    def reset_bc(self) -> None:
        _check(_ffi.hip().axbc_reset_bc(self.ctx))

    def set_bc(self, bc: np.ndarray) -> None:
        _check(_ffi.hip().axbc_set_bc(self.ctx, np.ascontiguousarray(bc, np.float64)))

    def set_checks(self, check) -> None:
        a = np.ascontiguousarray(check, dtype=np.uint32)
        _check(_ffi.hip().axbc_set_checks(self.ctx, a.ctypes.data, len(a)))

    def stop_test(self, c: int) -> bool:
        s = C.c_int(0)
        _check(_ffi.hip().axbc_stop_test(self.ctx, int(c), C.byref(s)))
        return bool(s.value)
