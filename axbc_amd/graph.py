"""Host-resident graphs in the CSC form the engine consumes.

Graph construction runs in libaxbc_host.so (C + OpenMP): the Kronecker
generator, mirror, normalisation, 1-degree reduction and CSC build, with the
semantics of AxBC's host code (src/axbc2d.c:333-907, 2544-2625).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _ffi

INITIATOR = (0.57, 0.19, 0.19, 0.05)


def _threads() -> int:
    n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


@dataclass
class Graph:
    """CSC graph: adjacency of u is row[col[u]:col[u+1]] (column = the second
    endpoint of the edge list, as build_csc, src/axbc2d.c:883-907)."""

    N: int
    col: np.ndarray  # uint32, N+1
    row: np.ndarray  # uint32, ned
    reach: np.ndarray  # uint32, N
    heuristic: int = 0
    generated: int = 0
    rem_ed: int = 0
    name: str = ""

    @property
    def ned(self) -> int:
        return int(self.row.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.col).astype(np.uint32)

    def footprint_bytes(self) -> int:
        return 4 * (self.N + 1) + 4 * self.ned + 4 * self.N

    @staticmethod
    def _from_c(g: _ffi.AxbcGraph, heuristic: int, name: str) -> "Graph":
        L = _ffi.host()
        n, ned = int(g.N), int(g.ned)
        col = np.ctypeslib.as_array(g.col, shape=(n + 1,)).copy()
        row = np.ctypeslib.as_array(g.row, shape=(max(ned, 1),))[:ned].copy()
        reach = np.ctypeslib.as_array(g.reach, shape=(n,)).copy()
        out = Graph(n, col, row, reach, heuristic, int(g.generated), int(g.rem_ed), name)
        L.axbc_graph_free(C.byref(g))
        return out

    @staticmethod
    def kronecker(scale: int, edgefactor: int = 16, seed: tuple[int, int] = (1, 2),
                  undirected: bool = True, heuristic: int = 0, threads: int | None = None) -> "Graph":
        """gen_graph (make_graph + mirror, src/axbc2d.c:518-527) followed by the
        1x1 preprocessing of main()."""
        L = _ffi.host()
        g = _ffi.AxbcGraph()
        rc = L.axbc_graph_kronecker(scale, edgefactor, seed[0], seed[1], int(undirected), heuristic,
                                    threads or _threads(), C.byref(g))
        if rc != 0:
            raise RuntimeError("graph construction failed")
        return Graph._from_c(g, heuristic, f"kron-s{scale}-ef{edgefactor}-seed{seed[0]}")

    @staticmethod
    def from_edges(N: int, edges: np.ndarray, undirected: bool = True, heuristic: int = 0,
                   threads: int | None = None, name: str = "edges") -> "Graph":
        """Build from an (m, 2) array of endpoints as read_graph would
        (mirror, then the same preprocessing)."""
        L = _ffi.host()
        e = np.ascontiguousarray(edges, dtype=np.uint64).reshape(-1, 2)
        if e.size and int(e.max()) >= N:
            raise ValueError("edge endpoint >= N")
        m = e.shape[0]
        keys = (e[:, 0] << np.uint64(32)) | e[:, 1]
        buf = L.malloc(max(8 * m, 8))
        C.memmove(buf, keys.ctypes.data, 8 * m)
        ned = C.c_uint64(m)
        ptr = L.axbc_mirror(buf, C.byref(ned), int(undirected))
        g = _ffi.AxbcGraph()
        rc = L.axbc_graph_from_edges(C.cast(ptr, C.c_void_p), ned.value, N, heuristic,
                                     threads or _threads(), C.byref(g))
        if rc != 0:
            raise RuntimeError("graph construction failed")
        out = Graph._from_c(g, heuristic, name)
        out.generated = m
        return out

    @staticmethod
    def from_keys_device(N: int, keys: np.ndarray, undirected: bool = True, heuristic: int = 0, device: int = 0,
                         name: str = "edges") -> "Graph":
        """The same preprocessing (mirror, normalise, 1-degree reduction,
        CSC) run on the device (axbc_ingest); keys are packed (u << 32 | v)."""
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        m = int(k.shape[0])
        col = np.empty(N + 1, np.uint32)
        row = np.empty(max(2 * m, 1), np.uint32)
        reach = np.empty(N, np.uint32)
        ned, rem = C.c_uint64(0), C.c_uint32(0)
        L = _ffi.hip()
        rc = L.axbc_ingest(device, k.ctypes.data, m, N, int(undirected), heuristic, col.ctypes.data, row.ctypes.data,
                           row.shape[0], reach.ctypes.data, C.byref(ned), C.byref(rem))
        if rc != 0:
            raise RuntimeError(L.axbc_last_error().decode())
        return Graph(N, col, row[:ned.value].copy(), reach, heuristic, m, int(rem.value), name)

    @staticmethod
    def kronecker_device(scale: int, edgefactor: int = 16, seed: tuple[int, int] = (1, 2),
                         undirected: bool = True, heuristic: int = 0, device: int = 0) -> "Graph":
        """Graph.kronecker built on the device: generation and preprocessing
        (axbc_kronecker_ingest), the same graph as the host pipeline's."""
        N, m = 1 << scale, (1 << scale) * edgefactor
        col = np.empty(N + 1, np.uint32)
        row = np.empty(max(2 * m, 1), np.uint32)
        reach = np.empty(N, np.uint32)
        ned, rem = C.c_uint64(0), C.c_uint32(0)
        ini = (C.c_double * 4)(*INITIATOR)
        L = _ffi.hip()
        rc = L.axbc_kronecker_ingest(device, scale, m, seed[0], seed[1], ini, int(undirected), heuristic,
                                     col.ctypes.data, row.ctypes.data, row.shape[0], reach.ctypes.data,
                                     C.byref(ned), C.byref(rem))
        if rc != 0:
            raise RuntimeError(L.axbc_last_error().decode())
        return Graph(N, col, row[:ned.value].copy(), reach, heuristic, m, int(rem.value),
                     f"kron-s{scale}-ef{edgefactor}-seed{seed[0]}")

    @staticmethod
    def read(path: str, N: int, undirected: bool = True, heuristic: int = 0) -> "Graph":
        """Text edge list ("u v" lines, '#' comments; src/axbc2d.c:356-445)."""
        L = _ffi.host()
        ned = C.c_uint64(0)
        err = C.create_string_buffer(512)
        ptr = L.axbc_read_edge_file(path.encode(), N, int(undirected), C.byref(ned), err, 512)
        if not ptr:
            raise ValueError(err.value.decode())
        g = _ffi.AxbcGraph()
        rc = L.axbc_graph_from_edges(C.cast(ptr, C.c_void_p), ned.value, N, heuristic, _threads(),
                                     C.byref(g))
        if rc != 0:
            raise RuntimeError("graph construction failed")
        return Graph._from_c(g, heuristic, os.path.basename(path))

    @staticmethod
    def lattice(side: int, p_delete: float = 0.03, seed: int = 3, heuristic: int = 0) -> "Graph":
        """side x side 4-neighbour lattice, edges deleted with p_delete, edges
        of the largest component kept, ids r*side+c (N = side^2)."""
        L = _ffi.host()
        ned = C.c_uint64(0)
        ptr = L.axbc_lattice_edges(side, p_delete, seed, C.byref(ned))
        ptr = L.axbc_mirror(C.cast(ptr, C.c_void_p), C.byref(ned), 1)
        g = _ffi.AxbcGraph()
        rc = L.axbc_graph_from_edges(C.cast(ptr, C.c_void_p), ned.value, side * side, heuristic,
                                     _threads(), C.byref(g))
        if rc != 0:
            raise RuntimeError("graph construction failed")
        return Graph._from_c(g, heuristic, f"lattice-{side}")


def kronecker_keys_device(scale: int, nedges: int, seed1: int, seed2: int, device: int = 0) -> np.ndarray:
    """The generator run on the device: packed keys (min << 32 | max) in the
    host generator's order (axbc_kronecker_generate)."""
    out = np.empty(max(nedges, 1), np.uint64)
    ini = (C.c_double * 4)(*INITIATOR)
    L = _ffi.hip()
    if L.axbc_kronecker_generate(device, scale, nedges, seed1, seed2, ini, out.ctypes.data) != 0:
        raise RuntimeError(L.axbc_last_error().decode())
    return out[:nedges]


def kronecker_edge_list(scale: int, nedges: int, seed1: int, seed2: int,
                        threads: int | None = None) -> np.ndarray:
    """Raw generator output as an (nedges, 2) int64 array ((min,max) edges,
    self-loops and duplicates kept)."""
    L = _ffi.host()
    ini = (C.c_double * 4)(*INITIATOR)
    ptr = L.axbc_kronecker_edges(scale, nedges, seed1, seed2, ini, threads or _threads())
    if not ptr:
        raise MemoryError("generator allocation failed")
    arr = np.ctypeslib.as_array(ptr, shape=(2 * nedges,)).astype(np.int64).reshape(-1, 2)
    L.free(C.cast(ptr, C.c_void_p))
    return arr
