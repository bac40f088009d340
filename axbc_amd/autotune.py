"""Autotuner for the batched sweep: picks, per graph, the number of roots
per batch (32, 64 or 128; 32 runs in the 64-wide kernels) and the split
threshold (edges per work item for high-degree vertices) by timing real
batches of the graph's own roots.  bc is saved and restored around the
trials, so tuning leaves the engine's result untouched.
"""
from __future__ import annotations

import time

import numpy as np

from . import _ffi
from .engine import Engine, _check

WIDTHS = (32, 64, 128)
CHUNKS = (128, 256, 512, 1024)


def _trial(e: Engine, roots: np.ndarray, width: int, chunk: int, reps: int) -> float:
    L = _ffi.hip()
    _check(L.axbc_set_batch_width(e.ctx, 128 if width == 128 else 64))
    _check(L.axbc_set_batch_chunk(e.ctx, chunk))
    p = np.full(width, 1.0 / e.N)
    e.batch(roots[:width], p[:width])  # warm-up (allocation, caches)
    e.sync()
    t = time.perf_counter()
    for r in range(reps):
        off = width * (r + 1)
        e.batch(roots[off:off + width], p[:width])
    e.sync()
    return (time.perf_counter() - t) / (reps * width)


def tune(e: Engine, roots: np.ndarray, reps: int = 2, widths=WIDTHS, chunks=CHUNKS) -> dict:
    """Time every (width, chunk) pair on `roots` (distinct vertices of
    positive degree, at least 128 * (reps + 1)); set the best on the engine
    and return {"width", "chunk", "s_per_root", "table"}."""
    roots = np.ascontiguousarray(roots, np.uint32)
    need = max(widths) * (reps + 1)
    if len(roots) < need:
        roots = np.resize(roots, need)
    saved = e.bc()
    table = {}
    for w in widths:
        for ch in chunks:
            table[(w, ch)] = _trial(e, roots, w, ch, reps)
    (w, ch), best = min(table.items(), key=lambda kv: kv[1])
    L = _ffi.hip()
    _check(L.axbc_set_batch_width(e.ctx, 128 if w == 128 else 64))
    _check(L.axbc_set_batch_chunk(e.ctx, ch))
    e.set_bc(saved)
    return {"width": w, "chunk": ch, "s_per_root": best,
            "table": {f"{k[0]}x{k[1]}": round(v * 1e3, 4) for k, v in table.items()}}
