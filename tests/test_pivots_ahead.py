"""Drawing the next batch's roots while a batch runs (PivotLoop.draw_ahead)
changes nothing a caller sees: after every step the roots, probabilities,
round/attempt/skip counts and the stop decision equal those of the loop that
draws each batch when it starts -- with steps of changing size, a stop inside
a batch, a stop before a batch, and degree-0 draws that consume attempts.
The device is replaced by a stand-in that applies a scripted number of roots
per batch."""
import ctypes as C

import numpy as np
import pytest

from axbc_amd import Graph
from axbc_amd import pivots as pv


class FakeLib:
    """Device calls of the batched loop: each batch applies all its roots
    unless `cut` says how many the stop test lets through for that batch."""

    def __init__(self, g, cut=None, stop_at=None):
        self.g, self.cut, self.stop_at = g, cut or {}, stop_at
        self.batches = 0
        self.tests = 0

    def axbc_set_checks(self, ctx, ptr, n):
        return 0

    def axbc_set_batch_width(self, ctx, w):
        return 0

    def axbc_hdn_build(self, ctx, pmax, sdn, s):
        n = self.g.N
        deg = np.diff(self.g.col).astype(np.uint64)
        sq = (deg * deg) & np.uint64(0xFFFFFFFF)
        pref = (np.cumsum(sq) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        self.runmax = np.maximum.accumulate(pref)
        if pmax:
            np.ctypeslib.as_array((C.c_uint32 * n).from_address(pmax))[:] = self.runmax
        if sdn:
            np.ctypeslib.as_array((C.c_uint32 * n).from_address(sdn))[:] = (deg + 1).astype(np.uint32)
        s._obj.value = int(pref[-1])
        return 0

    def axbc_draw_window(self, ctx, table, state, total, count, lr, v):
        # the windows' contract, restated: lrand48 values ahead of the state
        # and the first running-max entry >= r
        x = state
        m = (total + 1) & 0xFFFFFFFF
        for p in range(count):
            x = (0x5DEECE66D * x + 0xB) % (1 << 48)
            lr[p] = x >> 17
            r = int(lr[p]) % m if m else int(lr[p])
            v[p] = min(int(np.searchsorted(self.runmax, np.uint32(r), side="left")), self.g.N - 1)
        self.windows = getattr(self, "windows", 0) + 1
        return 0

    def axbc_stop_test(self, ctx, c, out):
        self.tests += 1
        out._obj.value = int(self.stop_at is not None and self.tests >= self.stop_at)
        return 0

    def axbc_batch(self, ctx, v, pp, n, c, ap, stats):
        k = self.cut.get(self.batches, n)
        self.batches += 1
        ap._obj.value = min(k, n)
        return 0


class FakeEngine:
    ctx = None

    def __init__(self, g):
        self.g = g

    def set_checks(self, check):
        pass

    def stop_test(self, c):
        s = C.c_int(0)
        pv._ffi.hip().axbc_stop_test(None, c, C.byref(s))
        return bool(s.value)


def run(monkeypatch, g, ahead, sizes, z, c=0, nbfs=400, cut=None, stop_at=None):
    fake = FakeLib(g, cut, stop_at)
    monkeypatch.setattr(pv._ffi, "hip", lambda: fake)
    loop = pv.PivotLoop(FakeEngine(g), nbfs, c=c, x=2, z=z, batch=64)
    loop.draw_ahead = ahead
    trace = []
    for m in sizes + [None] * 40:
        n = loop.step(m)
        r = loop.res
        trace.append((n, len(r.pivots), r.nrounds, r.skipped, r.attempts, r.stopped, loop.ui))
        if n == 0 and (r.stopped or loop.ui >= loop.nbfs):
            break
    return loop.res, trace


GRAPH = None


def graph():
    global GRAPH
    if GRAPH is None:
        # edgefactor 2 leaves many degree-0 vertices, whose draws consume
        # attempts without a round
        GRAPH = Graph.kronecker(10, 2, seed=(1, 2), heuristic=0)
    return GRAPH


@pytest.mark.parametrize("z", [0, 1])
@pytest.mark.parametrize("case", ["plain", "sizes", "cut", "stop"])
def test_draw_ahead_is_invisible(monkeypatch, z, case):
    g = graph()
    assert (np.diff(g.col) == 0).sum() > 100
    kw = {}
    sizes = []
    if case == "sizes":
        sizes = [64, 64, 32, 32, 64, 10, 64, 64]
    if case == "cut":
        kw = dict(c=1, cut={1: 17})
    if case == "stop":
        kw = dict(c=1, stop_at=3)
    nbfs = 400 if z == 0 else 200
    a_res, a_tr = run(monkeypatch, g, True, sizes, z, nbfs=nbfs, **kw)
    b_res, b_tr = run(monkeypatch, g, False, sizes, z, nbfs=nbfs, **kw)
    assert a_tr == b_tr
    assert a_res.pivots == b_res.pivots and a_res.probs == b_res.probs
    assert len(a_res.pivots) > 64
    if z == 0:
        assert a_res.skipped > 0
    if case in ("cut", "stop"):
        assert a_res.stopped


@pytest.mark.parametrize("window", [7, 37, 1024])
@pytest.mark.parametrize("ahead", [False, True])
def test_draw_windows_equal_per_draw_search(monkeypatch, window, ahead):
    """-z 1 draws taken from device windows (here the stand-in's) equal the
    per-draw search over the running maximum, whatever the window size and
    with or without drawing ahead -- windows cross batch boundaries and
    draw-ahead undos."""
    g = graph()
    monkeypatch.setattr(pv.PivotLoop, "WINDOW", window)
    sizes = [64, 64, 32, 10, 64]
    res, _ = run(monkeypatch, g, ahead, sizes, 1, nbfs=300)
    # the same loop restated with a search per draw
    fake = FakeLib(g)
    monkeypatch.setattr(pv._ffi, "hip", lambda: fake)
    ref = pv.PivotLoop(FakeEngine(g), 300, c=0, x=2, z=1, batch=64)

    def per_draw(table, total):
        r = ref.rng.uniform_int(total)
        return min(int(np.searchsorted(fake.runmax, np.uint32(r), side="left")), g.N - 1)

    ref._window_pick = per_draw
    ref.draw_ahead = False
    for m in sizes + [None] * 40:
        if ref.step(m) == 0 and ref.ui >= ref.nbfs:
            break
    assert res.pivots == ref.res.pivots and res.probs == ref.res.probs
    assert res.skipped == ref.res.skipped and len(res.pivots) > 100
