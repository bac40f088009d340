"""Batches of more than 128 roots through the bit-packed sweep (width 256,
batch mode 0 or 1): the roots from 128 on are swept in a second state on a
second stream beside the first 128, and the two halves are applied in draw
order.  The result must be the one of two batches run one after the other
(bc bit for bit, per-root counters, delta rows, levels, utilities), the
stop test must stop inside either half exactly where the one-root loop
stops, and list growth must still retry the whole batch."""
import ctypes as C

import numpy as np
import pytest

from axbc_amd import Engine, Graph, _ffi
from axbc_amd.pivots import PivotLoop
from oracle import oracle
from tests.parity import close
from tests.test_gpu_batch import delta_of, run_batch

pytestmark = pytest.mark.gpu


def test_require(require_gpu):
    pass


def _engine(g, width, mode=1):
    e = Engine(g)
    L = _ffi.hip()
    assert L.axbc_set_batch_mode(e.ctx, mode) == 0
    assert L.axbc_set_batch_width(e.ctx, width) == 0
    return e


def _utility(e, verts, wts, width):
    out = np.zeros(width, np.float64)
    assert _ffi.hip().axbc_batch_utility(e.ctx, verts.ctypes.data, wts.ctypes.data, len(verts), out.ctypes.data) == 0
    return out


def _roots(g, n, seed):
    deg = g.degrees()
    rng = np.random.default_rng(seed)
    return rng.choice(np.flatnonzero(deg > 0), size=n, replace=False).astype(np.uint32)


@pytest.mark.parametrize("nb,mode", [(256, 1), (200, 1), (129, 0)])
def test_pair_equals_two_batches(require_gpu, nb, mode):
    g = Graph.kronecker(12, 16, seed=(3, 5), heuristic=1)
    roots = _roots(g, nb, nb)
    probs = np.full(nb, 1.0 / g.N)
    verts = np.random.default_rng(1).choice(g.N, size=500, replace=False).astype(np.uint32)
    wts = np.random.default_rng(2).uniform(0.1, 2.0, size=500)
    print(f"[pair] N={g.N} nb={nb} mode={mode}", flush=True)
    with _engine(g, 128, mode) as seq:
        a0, s0 = run_batch(seq, roots[:128], probs[:128])
        d0 = {k: delta_of(seq, k) for k in (0, 77, 127)}
        l0 = seq.batch_levels(77)
        u0 = _utility(seq, verts, wts, 128)
        a1, s1 = run_batch(seq, roots[128:], probs[128:])
        d1 = {k: delta_of(seq, k) for k in (0, nb - 129)}
        l1 = seq.batch_levels(nb - 129)
        u1 = _utility(seq, verts, wts, 128)
        bc_seq = seq.bc()
    with _engine(g, 256, mode) as par:
        ap, st = run_batch(par, roots, probs)
        assert (a0, a1) == (128, nb - 128) and ap == nb
        assert st == s0 + s1
        assert np.array_equal(par.bc(), bc_seq), "bc differs from two batches in turn"
        for k, want in d0.items():
            assert np.array_equal(delta_of(par, k), want), f"delta of root {k}"
        for k, want in d1.items():
            assert np.array_equal(delta_of(par, 128 + k), want), f"delta of root {128 + k}"
        for got, want in zip(par.batch_levels(77), l0):
            assert np.array_equal(got, want)
        for got, want in zip(par.batch_levels(nb - 1), l1):
            assert np.array_equal(got, want)
        u = _utility(par, verts, wts, 256)
        assert np.array_equal(u[:128], u0) and np.array_equal(u[128:nb], u1[:nb - 128])
        # a batch of at most 128 roots afterwards runs in the first state
        more = _roots(g, 90, 99)
        a2, _ = run_batch(par, more, np.full(90, 1.0 / g.N))
        assert a2 == 90
        bc_ref = np.zeros(g.N)
        for v0 in np.concatenate([roots, more]):
            bc_ref = oracle.pivot(g, int(v0), 1.0 / g.N, bc_ref)[3]
        ok, err = close(par.bc(), bc_ref)
        assert ok, err


@pytest.mark.parametrize("half", [0, 1])
def test_pair_stop_inside_either_half(require_gpu, half):
    g = Graph.kronecker(10, 16, seed=(1, 2))
    deg = g.degrees()
    roots = _roots(g, 256, 3)
    probs = np.full(256, 1.0 / g.N)
    hub = int(np.argmax(deg))
    # the oracle sequence and the bc of the hub after each root
    bc_ref = np.zeros(g.N)
    seq = []
    for v0 in roots:
        bc_ref = oracle.pivot(g, int(v0), 1.0 / g.N, bc_ref)[3]
        seq.append(bc_ref[hub] / 2.0)
    # a threshold c N that the hub's bc crosses first inside the chosen half
    k0 = 60 if half == 0 else 190
    c = int(seq[k0] // g.N)
    first = next(k for k, x in enumerate(seq) if x >= c * g.N)
    assert c >= 1 and (first < 128) == (half == 0), (c, first)
    bc_want = np.zeros(g.N)
    for v0 in roots[:first + 1]:
        bc_want = oracle.pivot(g, int(v0), 1.0 / g.N, bc_want)[3]
    with _engine(g, 256) as e:
        e.set_checks(np.array([hub], np.uint32))
        applied, _ = run_batch(e, roots, probs, stop_c=c)
        assert applied == first + 1
        ok, err = close(e.bc(), bc_want)
        assert ok, err


def test_pair_list_growth(require_gpu):
    """256 roots spread along a path list most vertices at many distinct
    levels: both states' level lists outgrow their first capacity, and the
    batch is rerun from scratch with bc untouched until it fits."""
    n = 3000
    g = Graph.from_edges(n, np.array([(i, i + 1) for i in range(n - 1)]))
    roots = ((np.arange(256) * 11 + 7) % n).astype(np.uint32)
    assert len(set(roots.tolist())) == 256
    probs = np.full(256, 1.0 / n)
    with _engine(g, 256) as e:
        applied, _ = run_batch(e, roots, probs)
        assert applied == 256
        bc_ref = np.zeros(n)
        for k, v0 in enumerate(roots):
            lvl, sig, dl, bc_ref, _ = oracle.pivot(g, int(v0), 1.0 / n, bc_ref)
            if k in (0, 127, 128, 255):
                ok, err = close(delta_of(e, k), dl)
                assert ok, f"root #{k}: {err}"
        ok, err = close(e.bc(), bc_ref)
        assert ok, err


def test_pair_with_kernel_timing(require_gpu):
    """Per-kernel events (host launches on both streams) give the same bc as
    the level graphs."""
    g = Graph.kronecker(11, 16, seed=(7, 8), heuristic=1)
    roots = _roots(g, 250, 5)
    probs = np.full(250, 1.0 / g.N)
    out = []
    for timing in (0, 1):
        with _engine(g, 256) as e:
            assert _ffi.hip().axbc_kernel_timing(e.ctx, timing) == 0
            applied, _ = run_batch(e, roots, probs)
            assert applied == 250
            if timing:
                ms = C.c_double()
                n = C.c_uint64()
                b = C.c_uint64()
                L = _ffi.hip()
                L.axbc_kernel_stats.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                                C.POINTER(C.c_uint64)]
                assert L.axbc_kernel_stats(e.ctx, 3, C.byref(ms), C.byref(n), C.byref(b)) == 0
                assert n.value > 0 and ms.value > 0
            out.append(e.bc())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("z", [0, 1])
def test_pair_pivot_loop_equals_oracle_loop(require_gpu, z):
    """The adaptive loop in batches of 256 (static distributions): the same
    pivots, rounds and stop as the oracle's one-root loop."""
    g = Graph.kronecker(11, 16, seed=(3, 4), heuristic=1)
    nbfs = 700
    pv, pr, bc_ref, rs = oracle.run(g, nbfs, c=1, x=4, z=z)
    with _engine(g, 256, 0) as e:
        res = PivotLoop(e, nbfs, c=1, x=4, z=z, batch=256).run()
        assert res.pivots == [int(v) for v in pv]
        assert res.nrounds == rs["nrounds"] and res.stopped == rs["stopped"]
        ok, err = close(e.bc(), bc_ref)
        assert ok, err
