"""ctypes view of the CPU oracle (oracle/bin/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py as the checker, never by the engine.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "bin", "liboracle.so")
# a sanitizer build of the oracle (python -m axbc_amd.build --sanitize)
if os.environ.get("AXBC_SAN_DIR") and os.path.exists(os.path.join(os.environ["AXBC_SAN_DIR"], "liboracle.so")):
    LIB = os.path.join(os.environ["AXBC_SAN_DIR"], "liboracle.so")

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Stats(C.Structure):
    _fields_ = [("nvisited", C.c_uint64), ("nlevels", C.c_uint64), ("e_r", C.c_uint64),
                ("e_dag", C.c_uint64), ("all", C.c_uint64)]


class Cfg(C.Structure):
    _fields_ = [("nbfs", C.c_uint64), ("bader", C.c_uint32), ("approx_n", C.c_uint32),
                ("distribution", C.c_int), ("heuristic", C.c_int), ("random_seed_mode", C.c_int),
                ("seed", C.c_int64), ("sync_batch", C.c_uint32)]


class RunStats(C.Structure):
    _fields_ = [("nrounds", C.c_uint32), ("skipped", C.c_uint32), ("stopped", C.c_int),
                ("attempts", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB)
        L.oracle_pivot.restype = None
        L.oracle_pivot.argtypes = [C.c_uint32, u32p, u32p, u32p, C.c_int, C.c_uint32, C.c_double, i32p,
                                   u32p, f64p, f64p, C.POINTER(Stats)]
        L.oracle_run.restype = C.c_int
        L.oracle_pivots_omp.restype = C.c_uint64
        L.oracle_pivots_omp.argtypes = [C.c_uint32, u32p, u32p, u32p, C.c_int, u32p, f64p, C.c_uint32, C.c_int,
                                        f64p]
        L.oracle_run.argtypes = [C.c_uint32, u32p, u32p, u32p, C.POINTER(Cfg), u32p, f64p, f64p,
                                 C.POINTER(RunStats)]
        L.oracle_draws.restype = C.c_int
        L.oracle_draws.argtypes = [C.c_uint32, u32p, u32p, C.POINTER(Cfg), u32p, f64p, C.POINTER(RunStats)]
        L.oracle_lim_x_sample.restype = C.c_uint32
        L.oracle_lim_x_sample.argtypes = [C.c_uint64]
        L.oracle_hdn_sample.restype = C.c_uint32
        L.oracle_hdn_sample.argtypes = [C.c_uint32, u32p, u32p, C.c_uint32, C.POINTER(C.c_double)]
        L.oracle_fixed_weights.restype = C.c_uint64
        L.oracle_fixed_weights.argtypes = [C.c_uint32, f64p, u64p]
        L.oracle_weighted_sample.restype = C.c_uint32
        L.oracle_weighted_sample.argtypes = [C.c_uint32, u64p, C.c_uint64, C.c_long, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _arr(x, dt):
    return np.ascontiguousarray(x, dtype=dt)


def pivot(g, v0: int, p: float, bc: np.ndarray | None = None):
    """One oracle round on graph g (col/row/reach/heuristic/N attributes).
    Returns (lvl, sigma, delta, bc, stats); bc is accumulated in place when
    given."""
    N = g.N
    lvl = np.empty(N, np.int32)
    sig = np.empty(N, np.uint32)
    dl = np.empty(N, np.float64)
    if bc is None:
        bc = np.zeros(N, np.float64)
    st = Stats()
    row = _arr(g.row, np.uint32) if len(g.row) else np.zeros(1, np.uint32)
    lib().oracle_pivot(N, _arr(g.col, np.uint32), row, _arr(g.reach, np.uint32), int(g.heuristic),
                       int(v0), float(p), lvl, sig, dl, bc, C.byref(st))
    return lvl, sig, dl, bc, {k: int(getattr(st, k)) for k, _ in Stats._fields_}


def pivots_parallel(g, roots, probs, threads: int, bc: np.ndarray | None = None):
    """The rounds of the given roots over `threads` OpenMP threads (a
    throughput baseline for static distributions); returns (bc, summed E_r)."""
    N = g.N
    if bc is None:
        bc = np.zeros(N, np.float64)
    r = np.ascontiguousarray(roots, np.uint32)
    p = np.ascontiguousarray(probs, np.float64)
    row = _arr(g.row, np.uint32) if len(g.row) else np.zeros(1, np.uint32)
    er = lib().oracle_pivots_omp(N, _arr(g.col, np.uint32), row, _arr(g.reach, np.uint32), int(g.heuristic),
                                 r, p, len(r), int(threads), bc)
    return bc, int(er)


def run(g, nbfs: int, c: int = 1, x: int = 1, z: int = 0, seed: int = 4343543, sync_batch: int = 0):
    """Full pivot loop; returns (pivots, probs, bc, stats dict).  sync_batch
    > 0 freezes dynamic distributions for every group of that many rounds."""
    N = g.N
    cfg = Cfg(nbfs, c, x, z, int(g.heuristic), 0, seed, sync_batch)
    pv = np.zeros(max(1, min(nbfs, N)), np.uint32)
    pr = np.zeros(max(1, min(nbfs, N)), np.float64)
    bc = np.zeros(N, np.float64)
    rs = RunStats()
    row = _arr(g.row, np.uint32) if len(g.row) else np.zeros(1, np.uint32)
    rc = lib().oracle_run(N, _arr(g.col, np.uint32), row, _arr(g.reach, np.uint32), C.byref(cfg), pv, pr,
                          bc, C.byref(rs))
    if rc != 0:
        raise ValueError("unsupported configuration")
    n = rs.nrounds
    return pv[:n].copy(), pr[:n].copy(), bc, {"nrounds": n, "skipped": rs.skipped,
                                               "stopped": bool(rs.stopped), "attempts": rs.attempts}


def draws(g, nbfs: int, x: int = 1, z: int = 1, seed: int = 4343543):
    """The pivot sequence of run() for a static distribution (z = 0, 1, 2)
    with the stop test off, without running the rounds; returns (pivots,
    probs, stats dict)."""
    N = g.N
    cfg = Cfg(nbfs, 0, x, z, int(g.heuristic), 0, seed, 0)
    pv = np.zeros(max(1, min(nbfs, N)), np.uint32)
    pr = np.zeros(max(1, min(nbfs, N)), np.float64)
    rs = RunStats()
    row = _arr(g.row, np.uint32) if len(g.row) else np.zeros(1, np.uint32)
    rc = lib().oracle_draws(N, _arr(g.col, np.uint32), row, C.byref(cfg), pv, pr, C.byref(rs))
    if rc != 0:
        raise ValueError("draws need a static distribution (z = 0, 1, 2)")
    n = rs.nrounds
    return pv[:n].copy(), pr[:n].copy(), {"nrounds": n, "skipped": rs.skipped, "attempts": rs.attempts}

