"""Autotuner for the batched sweep: picks, per graph, the number of roots
per batch (the two widths of the bit-packed kernels, 64 and 128, and 256:
two 128-lane sweeps side by side; a batch of 32 is a partial batch of the
64-lane kernels, not a width of its own, and is tuned only when asked for),
the split threshold (edges per work item for high-degree vertices) and the
bottom-up switch ratio by timing real batches of the graph's own roots.
Width and threshold are searched jointly, the switch ratio afterwards at
the chosen pair.  bc is saved and restored around the trials, so tuning
leaves the engine's result untouched.
"""
from __future__ import annotations

import time

import numpy as np

from . import _ffi
from .engine import Engine, EngineError, _check

WIDTHS = (64, 128, 256)
CHUNKS = (128, 256, 512, 1024)
PULLS = (0, 16, 32, 64)


def _trial(e: Engine, roots: np.ndarray, width: int, chunk: int, reps: int, pull: int = 32) -> float:
    L = _ffi.hip()
    _check(L.axbc_set_batch_width(e.ctx, width if width >= 128 else 64))
    _check(L.axbc_set_batch_chunk(e.ctx, chunk))
    _check(L.axbc_set_batch_pull(e.ctx, pull))
    p = np.full(width, 1.0 / e.N)
    e.batch(roots[:width], p[:width])  # warm-up (allocation, caches)
    e.sync()
    t = time.perf_counter()
    for r in range(reps):
        off = width * (r + 1)
        e.batch(roots[off:off + width], p[:width])
    e.sync()
    return (time.perf_counter() - t) / (reps * width)


def tune(e: Engine, roots: np.ndarray, reps: int = 2, widths=WIDTHS, chunks=CHUNKS, pulls=PULLS) -> dict:
    """Time every (width, chunk) pair on `roots` (distinct vertices of
    positive degree, at least 128 * (reps + 1)), then every pull ratio at
    the best pair; set the best on the engine and return {"width", "chunk",
    "pull", "s_per_root", "table"} (table: ms per root)."""
    roots = np.ascontiguousarray(roots, np.uint32)
    need = max(widths) * (reps + 1)
    if len(roots) < need:
        roots = np.resize(roots, need)
    saved = e.bc()
    table = {}
    for w in widths:
        try:
            for ch in chunks:
                table[(w, ch, 32)] = _trial(e, roots, w, ch, reps, 32)
        except EngineError:
            # width 256 needs a second batch state; where the device memory
            # or the context's budget has no room for it, the width is out
            if w <= 128:
                raise
            for ch in chunks:
                table.pop((w, ch, 32), None)
            _check(_ffi.hip().axbc_set_batch_width(e.ctx, 128))
    (w, ch, _), _best = min(table.items(), key=lambda kv: kv[1])
    for pl in pulls:
        if (w, ch, pl) not in table:
            table[(w, ch, pl)] = _trial(e, roots, w, ch, reps, pl)
    (w, ch, pl), best = min(table.items(), key=lambda kv: kv[1])
    L = _ffi.hip()
    _check(L.axbc_set_batch_width(e.ctx, w if w >= 128 else 64))
    _check(L.axbc_set_batch_chunk(e.ctx, ch))
    _check(L.axbc_set_batch_pull(e.ctx, pl))
    e.set_bc(saved)
    return {"width": w, "chunk": ch, "pull": pl, "s_per_root": best,
            "table": {f"{k[0]}x{k[1]}x{k[2]}": round(v * 1e3, 4) for k, v in table.items()}}
