"""The pivot loop of AxBC's main() (src/axbc2d.c:2700-3156) over an Engine.

Same draw order as the reference (srand48(4343543), the -x check vertices,
the discarded initial uniform draw for -z 0, one lrand48 per attempt plus
the neighbour draw for -z 1), the same retry/skip rules (LIMxSAMPLE = N*N
as uint32, degree-0 roots consume an attempt), and the adaptive stop test
before every attempt.  The samplers' tables and the stop test run on the
device.  With batch > 0 the rounds of a static distribution are grouped
into bit-packed batches; the stop test is then evaluated exactly after
every root inside the batch, so the result equals the one-root loop.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _ffi
from .engine import Engine, _check

_libc = C.CDLL(None)
_libc.nrand48.restype = C.c_long
_libc.nrand48.argtypes = [C.POINTER(C.c_ushort * 3)]

DEFAULT_SEED = 4343543


class Rand48:
    """The lrand48 stream of srand48(seed), with its own state (nrand48), so
    that concurrent loops do not share libc's global generator."""

    def __init__(self, seed: int):
        s = seed & 0xFFFFFFFF
        self.x = (C.c_ushort * 3)(0x330E, s & 0xFFFF, (s >> 16) & 0xFFFF)
        self.pos = 0  # values drawn so far

    def lrand48(self) -> int:
        self.pos += 1
        return int(_libc.nrand48(C.byref(self.x)))

    def state48(self) -> int:
        """The 48-bit state X (the next value is of the advanced X)."""
        return self.x[0] | (self.x[1] << 16) | (self.x[2] << 32)

    def snapshot(self):
        return tuple(self.x), self.pos

    def restore(self, snap) -> None:
        x, self.pos = snap
        for j, v in enumerate(x):
            self.x[j] = v

    def uniform_int(self, mx: int, mn: int = 0) -> int:
        """lrand48() % (max - min + 1) + min in uint32 (src/axbc2d.c:2156-2164)."""
        m = (mx - mn + 1) & 0xFFFFFFFF
        r = self.lrand48()
        return ((r & 0xFFFFFFFF) % m + mn) & 0xFFFFFFFF if m else (r + mn) & 0xFFFFFFFF


@dataclass
class LoopResult:
    pivots: list = field(default_factory=list)
    probs: list = field(default_factory=list)
    nrounds: int = 0
    skipped: int = 0
    stopped: bool = False
    attempts: int = 0
    seconds: float = 0.0
    e_r: int = 0
    batch_seconds: list = field(default_factory=list)


class PivotLoop:
    """Stateful pivot loop: draws follow libc's lrand48 stream."""

    # batched -z 0/1 loops draw the next batch's roots while a batch runs
    draw_ahead = True

    def __init__(self, engine: Engine, nbfs: int, c: int = 1, x: int = 1, z: int = 0,
                 seed: int = DEFAULT_SEED, batch: int = 0):
        if z not in range(7):
            raise ValueError("distributions are 0..6")
        if batch and z not in (0, 1, 3, 4, 5, 6):
            raise ValueError("batches need -z 0, 1 (exact) or 3, 4, 5, 6 (batch-synchronous)")
        self.e, self.g = engine, engine.g
        self.N = self.g.N
        self.nbfs = min(nbfs, self.N)
        self.c, self.x, self.z, self.batch = c, x, z, batch
        self.col, self.row = self.g.col, self.g.row
        self.deg = np.diff(self.col)
        self.rng = Rand48(seed)
        uniform_int = self.rng.uniform_int
        self.checks = np.array([uniform_int(self.N) for _ in range(x)], np.uint32)
        engine.set_checks(self.checks)
        if batch:
            if batch not in (32, 64, 128, 256):
                raise ValueError("roots per batch: 32, 64, 128 or 256")
            _check(_ffi.hip().axbc_set_batch_width(engine.ctx, batch if batch >= 128 else 64))
        self.mix = z
        if z == 0:
            uniform_int(self.N)  # discarded initial draw (axbc2d.c:2740)
        if z == 6:
            self.mix = uniform_int(4)  # initial mixture component (axbc2d.c:2798)
        # the tables and their prefix stay on the device (draws are searched
        # there); the host keeps the per-vertex weights for the probabilities
        if z in (2, 6):
            self.lw = np.empty(self.N, np.uint64)
            t = C.c_uint64(0)
            _check(_ffi.hip().axbc_lcc_build(engine.ctx, None, self.lw.ctypes.data, None, C.byref(t)))
            self.ltotal = int(t.value)
        if z in (1, 6):
            self.sdn = np.empty(self.N, np.uint32)
            s = C.c_uint32(0)
            _check(_ffi.hip().axbc_hdn_build(engine.ctx, None, self.sdn.ctypes.data, C.byref(s)))
            self.sum_deg = int(s.value)
        self.lim = (self.N * self.N) & 0xFFFFFFFF
        self.vs = np.zeros(self.N, np.uint8)
        self.ui = 0
        self.res = LoopResult()
        self._marks = None  # vertices marked in vs while drawing ahead
        self._ahead = None  # roots drawn during the previous batch
        self._win = {}      # table -> (first generator position, lrand48 values, picked vertices)

    # -- sampling ---------------------------------------------------------
    WINDOW = 1024  # draws searched on the device at a time

    def _window_pick(self, table: int, total: int) -> int:
        """The vertex a first draw of table 1 (-z 1) or 2 (-z 2) picks at the
        generator's next position, searched on the device a window of
        positions at a time (axbc_draw_window); consumes the draw."""
        pos = self.rng.pos
        w = self._win.get(table)
        if w is None or not (w[0] <= pos < w[0] + len(w[1])):
            lr = np.empty(self.WINDOW, np.uint32)
            v = np.empty(self.WINDOW, np.uint32)
            _check(_ffi.hip().axbc_draw_window(self.e.ctx, table, self.rng.state48(), int(total), self.WINDOW, lr, v))
            w = (pos, lr, v)
            self._win[table] = w
        k = pos - w[0]
        if self.rng.lrand48() != int(w[1][k]):
            raise RuntimeError("device draw window out of step with the generator")
        return int(w[2][k])

    def _fixed_draw(self, total: int) -> int:
        """r = lrand48()/RAND_MAX * total (uniform_double), clamped."""
        rd = float(self.rng.lrand48()) / 2147483647.0 * float(total)
        return total if rd >= float(total) else int(rd)

    def _device_weight(self, kind: int, v: int) -> tuple[int, int]:
        L = _ffi.hip()
        total, w = C.c_uint64(0), C.c_uint64(0)
        _check(L.axbc_weights(self.e.ctx, kind, C.byref(total)))
        if v < self.N:
            _check(L.axbc_weights_at(self.e.ctx, v, C.byref(w)))
        return int(w.value), int(total.value)

    def _draw(self) -> tuple[int, float]:
        if self.z != 6:
            return self._draw_component(self.z)
        N = self.N
        v0, _ = self._draw_component(self.mix)
        # mixture probability (axbc2d.c:3030-3047), then the next component
        p = 1.0 / N
        first = self.res.nrounds < self.batch if self.batch else self.res.nrounds == 0
        if first:
            p = p + 1.0 / N
            p = p + 1.0 / N
        elif self.batch:
            # batch-synchronous: bc and delta weights frozen at the batch start
            inside = v0 < N
            p = p + (float(self._wb6[v0]) if inside else 0.0) / float(self._tb6)
            p = p + (float(self._wd6[v0]) if inside else 0.0) / float(self._td6)
        else:
            wb, tb = self._device_weight(3, v0)
            p = p + float(wb) / float(tb)
            wd, td = self._device_weight(4, v0)
            p = p + float(wd) / float(td)
        inside = v0 < N
        p = p + (float(self.sdn[v0]) if inside else 0.0) / float(self.sum_deg)
        p = p + (float(self.lw[v0]) if inside else 0.0) / float(self.ltotal)
        p = p / 5.0
        self.mix = self.rng.uniform_int(2)
        return v0, p

    def _draw_component(self, z: int) -> tuple[int, float]:
        N = self.N
        uniform_int = self.rng.uniform_int
        first = self.res.nrounds < self.batch if self.batch else self.res.nrounds == 0
        if z == 0 or (z in (3, 4, 5) and first):
            return uniform_int(N), 1.0 / N
        if z in (3, 4, 5) and self.batch:
            # batch-synchronous: the distribution frozen at the batch start
            r = self._fixed_draw(self._frozen_total)
            v, p = C.c_uint32(0), C.c_double(0)
            _check(_ffi.hip().axbc_weights_pick(self.e.ctx, r, C.byref(v), C.byref(p)))
            return int(v.value), float(p.value)
        if z == 2:
            u = self._window_pick(2, self.ltotal)
            return u, float(self.lw[u]) / float(self.ltotal)
        if z == 1:
            u = self._window_pick(1, self.sum_deg)
            d = int(self.deg[u])
            if d > 0:
                nb = 0 if d == 1 else uniform_int(d - 1)
                u = int(self.row[self.col[u] + nb])
            return u, float(self.sdn[u]) / float(self.sum_deg)
        total = C.c_uint64(0)
        _check(_ffi.hip().axbc_weights(self.e.ctx, z, C.byref(total)))
        r = self._fixed_draw(int(total.value))
        v, p = C.c_uint32(0), C.c_double(0)
        _check(_ffi.hip().axbc_weights_pick(self.e.ctx, r, C.byref(v), C.byref(p)))
        return int(v.value), float(p.value)

    def _next_root(self):
        """One attempt: (v0, p) or None for a consumed attempt."""
        bsm = 0
        while True:
            if bsm > self.lim:
                return None
            v0, p = self._draw()
            if v0 < self.N and self.vs[v0]:
                bsm += 1
                continue
            break
        if v0 < self.N and self._marks is not None:
            self._marks.append(v0)
        if v0 >= self.N or self.deg[v0] == 0:
            self.res.skipped += 1
            if v0 < self.N:
                self.vs[v0] = 1
            return None
        self.vs[v0] = 1
        return v0, p

    def sample_roots(self, n: int) -> np.ndarray:
        """The next n roots of the loop's draw sequence (distinct, of positive
        degree), drawn without running rounds; for static distributions
        (-z 0, 1, 2) the loop's own roots would follow them."""
        if self.z not in (0, 1, 2):
            raise ValueError("roots of a static distribution only")
        out = []
        while len(out) < n and self.ui < self.nbfs:
            r = self._next_root()
            self.ui += 1
            if r is not None:
                out.append(r[0])
        return np.array(out, np.uint32)

    def _collect(self, want: int):
        """Draw the next roots, up to `want` for a batch and one otherwise:
        (roots, attempt index after each, skipped count before each)."""
        res = self.res
        pend, pend_ui, pend_skip = [], [], []
        while self.ui < self.nbfs and len(pend) < want:
            skipped_before = res.skipped
            r = self._next_root()
            self.ui += 1
            if r is None:
                continue
            pend.append(r)
            pend_ui.append(self.ui)
            pend_skip.append(skipped_before)
            if not self.batch:
                break
        return pend, pend_ui, pend_skip

    # -- drawing ahead ------------------------------------------------------
    # Static exact distributions (-z 0, 1) do not depend on the rounds run,
    # so the next batch's roots are drawn on a host thread while the device
    # runs the current batch (the foreign call releases the interpreter
    # lock).  Between steps the attempt and skip counts show the rounds
    # taken so far, as without drawing ahead; the draws are undone
    # (generator state, sampled marks) when they are not used as drawn: the
    # loop stops, or the next step asks for another count.
    def _draw_ahead(self, want: int):
        """Start drawing the next `want` roots: (thread, box, want, state
        before).  The thread leaves the roots, the counts after them and the
        marks it set in box."""
        snap = (self.rng.snapshot(), self.ui, self.res.skipped)
        box = {}

        def work():
            self._marks = []
            try:
                box["drawn"] = self._collect(want)
                box["after"] = (self.ui, self.res.skipped)
            except BaseException as ex:  # raised again by the joiner
                box["error"] = ex
            box["marks"] = self._marks
            self._marks = None

        t = threading.Thread(target=work, daemon=True)
        t.start()
        return t, box, want, snap

    def _join_ahead(self):
        """Wait for the draws; the counts go back to the state before them."""
        t, box, _, (_, ui, skipped) = self._ahead
        t.join()
        self.ui, self.res.skipped = ui, skipped
        if "error" in box:
            self._undo_ahead()
            raise box["error"]

    def _take_ahead(self):
        """The roots drawn ahead, with the counts moved past them."""
        _, box, _, _ = self._ahead
        self._ahead = None
        self.ui, self.res.skipped = box["after"]
        return box["drawn"]

    def _undo_ahead(self):
        """Discard the roots drawn ahead and restore the sampler state."""
        if self._ahead is None:
            return
        t, box, _, (rs, ui, skipped) = self._ahead
        t.join()
        self._ahead = None
        self.rng.restore(rs)
        self.ui, self.res.skipped = ui, skipped
        for v in box.get("marks", ()):
            self.vs[v] = 0

    # -- running ----------------------------------------------------------
    def step(self, max_rounds: int | None = None) -> int:
        """Run rounds until max_rounds (a batch for batched loops) or the
        end; returns the rounds applied (0 when the loop is over)."""
        res, e = self.res, self.e
        want = max_rounds or (self.batch or 1)
        if self._ahead is not None and (self._ahead[2] != want or res.stopped):
            self._undo_ahead()
        if self._ahead is None and (res.stopped or self.ui >= self.nbfs):
            return 0
        if self.c > 0 and e.stop_test(self.c):
            self._undo_ahead()
            res.stopped = True
            return 0
        if self.batch and self.z in (3, 4, 5) and self.res.nrounds >= self.batch:
            t = C.c_uint64(0)
            _check(_ffi.hip().axbc_weights(e.ctx, self.z, C.byref(t)))
            self._frozen_total = int(t.value)
        if self.batch and self.z == 6 and self.res.nrounds >= self.batch:
            L = _ffi.hip()
            t = C.c_uint64(0)
            self._wb6 = np.empty(self.N, np.uint64)
            _check(L.axbc_weights(e.ctx, 3, C.byref(t)))
            _check(L.axbc_weights_copy(e.ctx, self._wb6.ctypes.data))
            self._tb6 = int(t.value)
            self._wd6 = np.empty(self.N, np.uint64)
            _check(L.axbc_weights(e.ctx, 4, C.byref(t)))
            _check(L.axbc_weights_copy(e.ctx, self._wd6.ctypes.data))
            self._td6 = int(t.value)
        if self._ahead is not None:
            pend, pend_ui, pend_skip = self._take_ahead()
        else:
            pend, pend_ui, pend_skip = self._collect(want)
        if not pend:
            return 0
        t = time.perf_counter()
        if not self.batch:
            v0, p = pend[0]
            st = e.pivot(v0, p, stats=True)
            applied = 1
            res.e_r += st["e_r"]
        else:
            L = _ffi.hip()
            v = np.array([a for a, _ in pend], np.uint32)
            pp = np.array([b for _, b in pend], np.float64)
            st = (_ffi.PivotStats * len(v))()
            ap = C.c_int(0)
            if self.draw_ahead and self.z in (0, 1) and self.ui < self.nbfs:
                self._ahead = self._draw_ahead(want)
            try:
                rc = L.axbc_batch(e.ctx, v, pp, len(v), self.c, C.byref(ap), C.cast(st, C.c_void_p))
            finally:
                if self._ahead is not None:
                    self._join_ahead()
            if rc != 0:
                self._undo_ahead()
                _check(rc)
            applied = ap.value
            res.e_r += sum(int(st[k].e_r) for k in range(applied))
            if applied < len(pend):  # stopped inside the batch
                self._undo_ahead()
                res.stopped = True
                res.skipped = pend_skip[applied - 1]
                self.ui = pend_ui[applied - 1]
        res.batch_seconds.append(time.perf_counter() - t)
        for k in range(applied):
            res.pivots.append(pend[k][0])
            res.probs.append(pend[k][1])
        res.nrounds += applied
        res.attempts = self.ui
        return applied

    def run(self) -> LoopResult:
        t = time.perf_counter()
        while self.step():
            pass
        self.e.sync()
        self.res.seconds = time.perf_counter() - t
        return self.res
