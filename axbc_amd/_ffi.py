"""ctypes bindings of the engine's two native libraries.

libaxbc_hip.so is the only compute path: when it is missing or no GPU is
visible, the engine raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class PivotStats(C.Structure):
    _fields_ = [("nvisited", C.c_uint64), ("nlevels", C.c_uint64), ("e_r", C.c_uint64),
                ("e_dag", C.c_uint64), ("all", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class AxbcGraph(C.Structure):
    _fields_ = [("N", C.c_uint64), ("ned", C.c_uint64), ("generated", C.c_uint64),
                ("rem_ed", C.c_uint32), ("col", C.POINTER(C.c_uint32)),
                ("row", C.POINTER(C.c_uint32)), ("reach", C.POINTER(C.c_uint32))]


_hip = None
_host = None


def _load(name: str) -> C.CDLL:
    # AXBC_SAN_DIR: a directory of sanitizer builds of the host libraries
    # (python -m axbc_amd.build --sanitize), used where it holds the file
    san = os.environ.get("AXBC_SAN_DIR")
    path = os.path.join(san, name) if san and os.path.exists(os.path.join(san, name)) else os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m axbc_amd.build` "
                           "(the engine has no fallback path)")
    return C.CDLL(path)


def hip() -> C.CDLL:
    """libaxbc_hip.so with argument types declared (include/axbc.h)."""
    global _hip
    if _hip is not None:
        return _hip
    L = _load("libaxbc_hip.so")
    ctxp = C.c_void_p
    sig = {
        "axbc_last_error": (C.c_char_p, []),
        "axbc_device_count": (C.c_int, []),
        "axbc_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.c_uint64, u32p, u32p, C.c_void_p, C.c_int,
                                      C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
        "axbc_ctx_destroy": (C.c_int, [ctxp]),
        "axbc_pivot": (C.c_int, [ctxp, C.c_uint32, C.c_double, C.c_void_p]),
        "axbc_sync": (C.c_int, [ctxp]),
        "axbc_reset_bc": (C.c_int, [ctxp]),
        "axbc_get_bc": (C.c_int, [ctxp, f64p]),
        "axbc_get_delta": (C.c_int, [ctxp, f64p]),
        "axbc_get_deg": (C.c_int, [ctxp, u32p]),
        "axbc_get_lvl": (C.c_int, [ctxp, i32p]),
        "axbc_get_sigma": (C.c_int, [ctxp, u32p]),
        "axbc_set_bc": (C.c_int, [ctxp, f64p]),
        "axbc_set_checks": (C.c_int, [ctxp, C.c_void_p, C.c_uint32]),
        "axbc_stop_test": (C.c_int, [ctxp, C.c_uint32, C.POINTER(C.c_int)]),
        "axbc_begin_round": (C.c_int, [ctxp]),
        "axbc_set_root": (C.c_int, [ctxp, C.c_uint32]),
        "axbc_forward_level": (C.c_int, [ctxp, C.c_int, C.POINTER(C.c_uint32)]),
        "axbc_backward_depth": (C.c_int, [ctxp, C.c_int]),
        "axbc_reach_sum": (C.c_int, [ctxp, C.POINTER(C.c_uint32)]),
        "axbc_update_bc": (C.c_int, [ctxp, C.c_uint32, C.c_double, C.c_uint64]),
        "axbc_last_stats": (C.c_int, [ctxp, C.POINTER(PivotStats)]),
        "axbc_last_timing": (C.c_int, [ctxp, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]),
    }
    sig.update({
        "axbc_batch": (C.c_int, [ctxp, u32p, f64p, C.c_int, C.c_uint32, C.POINTER(C.c_int), C.c_void_p]),
        "axbc_batch_width": (C.c_int, [ctxp]),
        "axbc_set_batch_width": (C.c_int, [ctxp, C.c_int]),
        "axbc_set_batch_chunk": (C.c_int, [ctxp, C.c_uint32]),
        "axbc_set_batch_pull": (C.c_int, [ctxp, C.c_uint32]),
        "axbc_batch_utility": (C.c_int, [ctxp, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
        "axbc_weights_copy": (C.c_int, [ctxp, C.c_void_p]),
        "axbc_ingest": (C.c_int, [C.c_int, C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_void_p,
                                 C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint32)]),
        "axbc_kronecker_generate": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64,
                                              C.POINTER(C.c_double), C.c_void_p]),
        "axbc_kronecker_ingest": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64,
                                            C.POINTER(C.c_double), C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_uint64, C.c_void_p, C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_uint32)]),
        "axbc_set_batch_graphs": (C.c_int, [ctxp, C.c_int]),
        "axbc_set_batch_persistent": (C.c_int, [ctxp, C.c_int, C.c_double]),
        "axbc_set_batch_mode": (C.c_int, [ctxp, C.c_int]),
        "axbc_set_plain_pull": (C.c_int, [ctxp, C.c_uint32]),
        "axbc_set_memory_budget": (C.c_int, [ctxp, C.c_size_t]),
        "axbc_device_bytes": (C.c_int, [ctxp, C.POINTER(C.c_size_t)]),
        "axbc_draw_window": (C.c_int, [ctxp, C.c_int, C.c_uint64, C.c_uint64, C.c_uint32, u32p, u32p]),
        "axbc_batch_state_bytes": (C.c_int, [C.c_uint64, C.c_uint64, u32p, C.c_int, C.POINTER(C.c_size_t)]),
        "axbc_batch_delta": (C.c_int, [ctxp, C.c_int, f64p]),
        "axbc_batch_levels": (C.c_int, [ctxp, C.c_int, i32p, u32p]),
        "axbc_hdn_build": (C.c_int, [ctxp, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]),
        "axbc_hdn_pick": (C.c_int, [ctxp, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "axbc_weights": (C.c_int, [ctxp, C.c_int, C.POINTER(C.c_uint64)]),
        "axbc_weights_pick": (C.c_int, [ctxp, C.c_uint64, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
        "axbc_weights_at": (C.c_int, [ctxp, C.c_uint32, C.POINTER(C.c_uint64)]),
        "axbc_lcc_build": (C.c_int, [ctxp, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
        "axbc_set_reach": (C.c_int, [ctxp, C.c_void_p]),
        "axbc_kernel_timing": (C.c_int, [ctxp, C.c_int]),
        "axbc_kernel_timing_kinds": (C.c_int, [ctxp, C.c_uint32]),
        "axbc_kernel_stats": (C.c_int, [ctxp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64)]),
        "axbc_copy_kernel": (C.c_int, [ctxp, C.c_uint64, C.c_int]),
        "axbc_read_kernel": (C.c_int, [ctxp, C.c_uint64, C.c_int]),
        "axbc_gather_kernel": (C.c_int, [ctxp, C.c_uint64, C.c_int]),
        "axbc_kernel_busy": (C.c_int, [ctxp, C.c_int, C.POINTER(C.c_double)]),
    })
    optional = {}
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    for name, (res, args) in optional.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
    _hip = L
    return L


def host() -> C.CDLL:
    """libaxbc_host.so (graph tooling)."""
    global _host
    if _host is not None:
        return _host
    L = _load("libaxbc_host.so")
    L.axbc_kronecker_edges.restype = C.POINTER(C.c_uint64)
    L.axbc_kronecker_edges.argtypes = [C.c_int, C.c_int64, C.c_uint64, C.c_uint64,
                                       C.POINTER(C.c_double), C.c_int]
    L.axbc_graph_kronecker.restype = C.c_int
    L.axbc_graph_kronecker.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                       C.c_int, C.POINTER(AxbcGraph)]
    L.axbc_graph_from_edges.restype = C.c_int
    L.axbc_graph_from_edges.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                        C.POINTER(AxbcGraph)]
    L.axbc_graph_free.restype = None
    L.axbc_graph_free.argtypes = [C.POINTER(AxbcGraph)]
    L.axbc_lattice_edges.restype = C.POINTER(C.c_uint64)
    L.axbc_lattice_edges.argtypes = [C.c_uint32, C.c_double, C.c_uint64, C.POINTER(C.c_uint64)]
    L.axbc_mirror.restype = C.POINTER(C.c_uint64)
    L.axbc_mirror.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.c_int]
    L.axbc_read_edge_file.restype = C.POINTER(C.c_uint64)
    L.axbc_read_edge_file.argtypes = [C.c_char_p, C.c_uint64, C.c_int, C.POINTER(C.c_uint64),
                                      C.c_char_p, C.c_size_t]
    L.axbc_norm_graph.restype = C.c_uint64
    L.axbc_norm_graph.argtypes = [C.c_void_p, C.c_uint64, C.c_int]
    L.axbc_degree_reduction.restype = C.c_uint64
    L.axbc_degree_reduction.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, u32p, C.POINTER(C.c_uint32),
                                        C.c_int]
    L.axbc_sort_edges.restype = None
    L.axbc_sort_edges.argtypes = [C.c_void_p, C.c_uint64, C.c_int]
    L.axbc_mrg_seed_from_user.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p]
    L.axbc_mrg_skip.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64]
    L.axbc_mrg_next_uint.restype = C.c_uint32
    L.axbc_mrg_next_uint.argtypes = [C.c_void_p]
    L.axbc_mrg_next_double.restype = C.c_double
    L.axbc_mrg_next_double.argtypes = [C.c_void_p]
    L.axbc_btrd_binomial.restype = C.c_int64
    L.axbc_btrd_binomial.argtypes = [C.c_int64, C.c_double, C.c_void_p]
    L.free = C.CDLL(None).free
    L.free.argtypes = [C.c_void_p]
    L.malloc = C.CDLL(None).malloc
    L.malloc.restype = C.c_void_p
    L.malloc.argtypes = [C.c_size_t]
    _host = L
    return L
