"""Whole pivot loop (sampler + stop test + rounds) on the device against
the oracle's loop: same pivot sequence, same round/skip counts, same stop
decision, bc within 1e-9 relative -- one root per round and batched."""
import numpy as np
import pytest

from axbc_amd import Engine, Graph
from axbc_amd.pivots import PivotLoop
from oracle import oracle
from tests.parity import close

pytestmark = pytest.mark.gpu


def test_require(require_gpu):
    pass


@pytest.mark.parametrize("z,c,x,batch,heur", [
    (0, 0, 1, 0, 0), (0, 1, 1, 0, 1), (1, 0, 4, 0, 1), (3, 0, 1, 0, 0), (4, 0, 1, 0, 1), (5, 0, 1, 0, 0),
    (0, 0, 1, 64, 0), (1, 1, 8, 64, 1), (1, 0, 1, 128, 1), (0, 2, 3, 64, 0),
    (2, 0, 1, 0, 0), (6, 0, 2, 0, 1), (4, 0, 1, 64, 0), (3, 1, 2, 32, 1), (5, 0, 1, 128, 0),
    (6, 0, 2, 64, 1), (6, 1, 3, 128, 0),
])
def test_loop_matches_oracle(require_gpu, z, c, x, batch, heur):
    g = Graph.kronecker(11, 16, seed=(1, 2), heuristic=heur)
    nbfs = 300
    print(f"[loop z{z} c{c} x{x} B{batch} H{heur}]", flush=True)
    sync = batch if z in (3, 4, 5, 6) else 0
    pv, pr, bc_ref, rs = oracle.run(g, nbfs, c=c, x=x, z=z, sync_batch=sync)
    with Engine(g) as e:
        res = PivotLoop(e, nbfs, c=c, x=x, z=z, batch=batch).run()
        assert res.pivots == [int(v) for v in pv]
        # probabilities of static distributions are exact; dynamic ones are
        # ratios of fixed-point weights of the FP64 bc/delta, equal within
        # the FP64 tolerance
        np.testing.assert_allclose(np.array(res.probs), pr, rtol=0 if z in (0, 1, 2) else 1e-9)
        assert res.nrounds == rs["nrounds"] and res.skipped == rs["skipped"]
        assert res.stopped == rs["stopped"]
        ok, err = close(e.bc(), bc_ref)
        assert ok, err


def _looped():
    # the first sorted edge survives normalisation even as a loop (0, 0):
    # the LCC then takes the general (per-direction) path
    rng = np.random.default_rng(3)
    e = rng.integers(0, 300, size=(3000, 2))
    e = np.concatenate([[[0, 0]], e])
    return Graph.from_edges(300, e, heuristic=0)


def test_lcc_matches_oracle(require_gpu):
    """Symmetric loop-free graphs intersect every undirected edge once (with
    hub items above 4096 edges at SCALE 16); directed graphs and a graph
    with a loop take the per-direction kernels."""
    import ctypes as C
    from axbc_amd import _ffi
    from axbc_amd import kronecker_edge_list
    # a SCALE-12 edge list among 2^21 ids: from 2^21 vertices on, owners of
    # 256..4096 edges probe an LDS hash instead of a bitmap
    sparse = Graph.from_edges(1 << 21, kronecker_edge_list(12, 1 << 16, 5, 6), heuristic=1)
    graphs = (Graph.kronecker(12, 16, seed=(1, 2)), Graph.kronecker(11, 32, seed=(7, 8), undirected=False),
              Graph.kronecker(16, 16, seed=(2, 3), heuristic=1), _looped(), sparse)
    assert int(np.diff(graphs[2].col).max()) > 4096
    assert graphs[3].row[graphs[3].col[0]] == 0
    sd = np.diff(sparse.col)
    assert ((sd >= 256) & (sd <= 4096)).any()
    for g in graphs:
        print(f"[lcc] N={g.N} ned={g.ned}", flush=True)
        lcc_ref = np.zeros(g.N, np.float32)
        L = oracle.lib()
        L.oracle_lcc.argtypes = [C.c_uint32, oracle.u32p, oracle.u32p,
                                 np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")]
        L.oracle_lcc(g.N, g.col, g.row, lcc_ref)
        with Engine(g) as e:
            lcc = np.empty(g.N, np.float32)
            lw = np.empty(g.N, np.uint64)
            t = C.c_uint64(0)
            assert _ffi.hip().axbc_lcc_build(e.ctx, lcc.ctypes.data, lw.ctypes.data, None, C.byref(t)) == 0
        np.testing.assert_array_equal(lcc, lcc_ref)
        W = np.zeros(g.N, np.uint64)
        total = L.oracle_fixed_weights(g.N, lcc_ref.astype(np.float64), W)
        np.testing.assert_array_equal(lw, W)
        assert t.value == total


def test_autotuned_loop_matches_oracle(require_gpu):
    """The autotuner leaves bc untouched, and the loop it configures (width,
    split threshold, pull ratio) still gives the oracle's loop."""
    from axbc_amd.autotune import tune
    g = Graph.kronecker(11, 16, seed=(3, 4), heuristic=1)
    nbfs = 200
    pv, pr, bc_ref, rs = oracle.run(g, nbfs, c=1, x=4, z=1)
    with Engine(g) as e:
        e.pivot(int(np.argmax(g.degrees())), 0.5)
        before = e.bc()
        rng = np.random.default_rng(5)
        sample = rng.choice(np.flatnonzero(g.degrees() > 0), 384, replace=False)
        t = tune(e, sample, reps=1, widths=(32, 64, 128, 256), chunks=(64, 256), pulls=(0, 32, 65536))
        print(f"[autotune] {t['width']} {t['chunk']} {t['pull']}", flush=True)
        assert t["width"] in (32, 64, 128, 256) and t["chunk"] in (64, 256) and t["pull"] in (0, 32, 65536)
        np.testing.assert_array_equal(e.bc(), before)
        e.reset_bc()
        res = PivotLoop(e, nbfs, c=1, x=4, z=1, batch=t["width"]).run()
        assert res.pivots == [int(v) for v in pv]
        assert res.nrounds == rs["nrounds"] and res.stopped == rs["stopped"]
        ok, err = close(e.bc(), bc_ref)
        assert ok, err


def test_hdn_pick_over_wrapped_prefix(require_gpu):
    """The device's high-degree-neighbourhood search on a graph whose uint32
    (LOCINT)pow(deg, 2) prefix wraps (hubs of degree 50000, 40000, 70000 and
    62000): for draws across the wrap the first vertex whose prefix reaches
    r, against Python integers mod 2^32 and a linear first match
    (src/axbc2d.c:2188-2198, 2895-2928)."""
    import ctypes as C
    from axbc_amd import _ffi
    from tests.test_oracle_pins import hub_graph, hdn_expect
    g = hub_graph()
    deg, prefix = hdn_expect(g)
    pre = np.array(prefix, np.uint64)
    S = prefix[-1]
    runmax = np.maximum.accumulate(pre)
    L = _ffi.hip()
    with Engine(g) as e:
        pmax = np.empty(g.N, np.uint32)
        sdn = np.empty(g.N, np.uint32)
        s = C.c_uint32(0)
        assert L.axbc_hdn_build(e.ctx, pmax.ctypes.data, sdn.ctypes.data, C.byref(s)) == 0
        assert s.value == S
        assert np.array_equal(pmax.astype(np.uint64), runmax)
        rng = np.random.default_rng(9)
        rs = [0, 1, 2500000000, 2500000001, 4100000000, 4100000001, 4254065408, 4254065409, S, S + 1]
        rs += [int(x) for x in rng.integers(0, S + 1, size=200)]
        for r in rs:
            hits = np.flatnonzero(pre >= r)
            want = int(hits[0]) if hits.size else g.N - 1
            u, su = C.c_uint32(0), C.c_uint32(0)
            assert L.axbc_hdn_pick(e.ctx, r, C.byref(u), C.byref(su)) == 0
            assert u.value == want, f"r = {r}: device {u.value}, first match {want}"
            want_sdn = sum(deg[w] for w in g.row[g.col[want]:g.col[want + 1]]) & 0xFFFFFFFF
            assert su.value == want_sdn


def test_draw_windows_on_the_device(require_gpu):
    """axbc_draw_window: the lrand48 values ahead of a generator state equal
    the generator's own (Rand48, itself pinned against libc), and the vertex
    at each position is the first match of that draw -- over the wrapped
    deg^2 prefix of a hub graph (table 1, linear first match in Python
    integers) and over the LCC weight prefix of a Kronecker graph (table 2,
    uniform_double restated)."""
    import ctypes as C
    from axbc_amd import _ffi
    from axbc_amd.pivots import Rand48
    from tests.test_oracle_pins import hub_graph, hdn_expect
    L = _ffi.hip()
    g = hub_graph()
    deg, prefix = hdn_expect(g)
    pre = np.array(prefix, np.uint64)
    S = prefix[-1]
    with Engine(g) as e:
        s = C.c_uint32(0)
        sdn = np.empty(g.N, np.uint32)
        assert L.axbc_hdn_build(e.ctx, None, sdn.ctypes.data, C.byref(s)) == 0
        for seed, skip, count in ((4343543, 0, 1000), (77, 12345, 4096), (1, 3, 1)):
            rng = Rand48(seed)
            for _ in range(skip):
                rng.lrand48()
            lr = np.empty(count, np.uint32)
            v = np.empty(count, np.uint32)
            assert L.axbc_draw_window(e.ctx, 1, rng.state48(), S, count, lr, v) == 0
            for p in range(count):
                x = rng.lrand48()
                assert lr[p] == x, (seed, p)
                r = x % ((S + 1) & 0xFFFFFFFF)
                hits = np.flatnonzero(pre >= r)
                assert v[p] == (int(hits[0]) if hits.size else g.N - 1), (seed, p, r)
        bad = np.empty(1, np.uint32)
        assert L.axbc_draw_window(e.ctx, 2, 0, 1, 1, bad, bad) != 0  # LCC table not built
    g = Graph.kronecker(12, 16, seed=(3, 4), heuristic=1)
    with Engine(g) as e:
        lw = np.empty(g.N, np.uint64)
        lp = np.empty(g.N, np.uint64)
        t = C.c_uint64(0)
        assert L.axbc_lcc_build(e.ctx, None, lw.ctypes.data, lp.ctypes.data, C.byref(t)) == 0
        total = int(t.value)
        assert total > 0 and np.array_equal(np.cumsum(lw, dtype=np.uint64), lp)
        rng = Rand48(4343543)
        count = 3000
        lr = np.empty(count, np.uint32)
        v = np.empty(count, np.uint32)
        assert L.axbc_draw_window(e.ctx, 2, rng.state48(), total, count, lr, v) == 0
        for p in range(count):
            x = rng.lrand48()
            assert lr[p] == x
            rd = float(x) / 2147483647.0 * float(total)
            r = total if rd >= float(total) else int(rd)
            u = int(np.searchsorted(lp, np.uint64(r), side="left"))
            assert v[p] == min(u, g.N - 1), (p, r)
