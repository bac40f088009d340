"""The workgroup sweep (one workgroup per root, axbc_set_batch_mode 2)
against the CPU oracle: per-root levels and path counts bit-exact, delta and
bc within 1e-9, the per-root counters equal; bit-identical to the bit-packed
sweep on the same batches; batches of both kinds in turn on one engine; the
automatic choice on a deep graph; level-offset growth past 16384 levels."""
import ctypes as C

import numpy as np
import pytest

from axbc_amd import Engine, Graph, _ffi
from oracle import oracle
from tests.parity import close
from tests.test_gpu_batch import check_batch, delta_of, run_batch

pytestmark = pytest.mark.gpu


def test_require(require_gpu):
    pass


def _mode(e, mode):
    assert _ffi.hip().axbc_set_batch_mode(e.ctx, mode) == 0, _ffi.hip().axbc_last_error()


@pytest.mark.parametrize("width", [64, 128])
@pytest.mark.parametrize("scale,seed,undirected,heur", [
    (10, (1, 2), True, 0),
    (11, (23, 24), True, 1),
    (10, (3, 4), False, 0),
])
def test_kronecker_batches_workgroup_sweep(require_gpu, width, scale, seed, undirected, heur):
    g = Graph.kronecker(scale, 16, seed=seed, undirected=undirected, heuristic=heur)
    deg = g.degrees()
    rng = np.random.default_rng(scale + width)
    cand = np.flatnonzero(deg > 0)
    roots = rng.choice(cand, size=min(len(cand), width + width // 2), replace=False)
    check_batch(g, roots, width, f"wg kron s{scale} H{heur} U{int(not undirected)}", batches=2, mode=2)


def test_star_path_and_single_level_roots(require_gpu):
    # a hub of 5000 edges plus a 2000-level path; the star centre's round
    # has a single level (its own delta is accumulated: the do-while)
    n = 7000
    e = [(0, i) for i in range(1, 5000)] + [(i, i + 1) for i in range(4999, n - 1)]
    g = Graph.from_edges(n, np.array(e))
    check_batch(g, np.array([0, 1, 2, 4999, 6999, 5500, 3], np.uint32), 64, "wg star+path", mode=2)
    star = Graph.from_edges(40, np.array([(0, i) for i in range(1, 40)]))
    check_batch(star, np.array([0, 5, 7], np.uint32), 64, "wg star", mode=2)


@pytest.mark.parametrize("width", [64, 128])
def test_sigma_wraps_to_zero(require_gpu, width):
    from tests.test_gpu_plain import layered
    g = layered(256, 6)
    roots = np.array([0, 1 + 5 * 256 + 7, 1 + 2 * 256 + 3, 1 + 4 * 256, 5], np.uint32)
    check_batch(g, roots, width, "wg sigma wraps to 0", mode=2)


def test_level_offsets_grow(require_gpu):
    """A path of 40000 vertices: roots at its ends run 39999 levels, past
    the first level-offset capacity (16384), so the batch runs again with a
    larger one; a root in the middle runs 20000."""
    n = 40000
    g = Graph.from_edges(n, np.array([(i, i + 1) for i in range(n - 1)]))
    check_batch(g, np.array([0, n - 1, n // 2, 17], np.uint32), 64, "wg path 40000", mode=2)


@pytest.mark.parametrize("width", [64, 128])
def test_bit_identical_to_packed_sweep(require_gpu, width):
    """The same batches through both sweeps give the same bits: bc, every
    root's delta, levels and path counts, and the counters.  The bit-packed
    sweep runs with a chunk above every list here: split vertices add their
    dependency terms in chunk partials (another association of the same
    sum, equal within 1e-9 only)."""
    g = Graph.kronecker(12, 16, seed=(5, 6), heuristic=1)
    deg = g.degrees()
    roots = np.random.default_rng(width).choice(np.flatnonzero(deg > 0), size=2 * width, replace=False)
    probs = np.full(width, 1.0 / g.N)
    out = {}
    for mode in (1, 2):
        with Engine(g) as e:
            assert _ffi.hip().axbc_set_batch_width(e.ctx, width) == 0
            assert _ffi.hip().axbc_set_batch_chunk(e.ctx, 1 << 20) == 0
            _mode(e, mode)
            got = []
            for b0 in (0, width):
                _, stats = run_batch(e, roots[b0:b0 + width], probs)
                got.append((stats, [delta_of(e, k) for k in (0, width // 2, width - 1)],
                            [e.batch_levels(k) for k in (0, width - 1)]))
            out[mode] = (e.bc(), got)
    np.testing.assert_array_equal(out[1][0], out[2][0])
    for (s1, d1, l1), (s2, d2, l2) in zip(out[1][1], out[2][1]):
        assert s1 == s2
        for a, b in zip(d1, d2):
            np.testing.assert_array_equal(a, b)
        for (a1, a2), (b1, b2) in zip(l1, l2):
            np.testing.assert_array_equal(a1, b1)
            np.testing.assert_array_equal(a2, b2)


def test_sweeps_in_turn(require_gpu):
    """Batches alternate between the sweeps on one engine over two disjoint
    graphs (roots in A, then in B, ...), so that a workgroup sweep follows a
    bit-packed one that reached vertices it does not reach, and the other
    way round; every delta, level and counter and the bc against the
    oracle."""
    from axbc_amd import kronecker_edge_list
    ea = kronecker_edge_list(11, 16 << 11, 5, 6)
    eb = kronecker_edge_list(11, 16 << 11, 7, 8) + (1 << 11)
    g = Graph.from_edges(1 << 12, np.concatenate([ea, eb]), heuristic=1)
    deg = np.diff(g.col)
    rng = np.random.default_rng(13)
    ra = rng.choice(np.flatnonzero(deg[: 1 << 11] > 0), size=3 * 40, replace=False)
    rb = rng.choice(np.flatnonzero(deg[1 << 11:] > 0), size=2 * 40, replace=False) + (1 << 11)
    batches = [ra[:40], rb[:40], ra[40:80], rb[40:], ra[80:]]
    modes = [1, 2, 1, 2, 1]
    bc_ref = np.zeros(g.N)
    with Engine(g) as e:
        for rts, mode in zip(batches, modes):
            _mode(e, mode)
            applied, stats = run_batch(e, rts, np.full(len(rts), 1.0 / g.N))
            assert applied == len(rts)
            for k, v0 in enumerate(rts):
                lvl, sig, dl, bc_ref, ost = oracle.pivot(g, int(v0), 1.0 / g.N, bc_ref)
                ok, err = close(delta_of(e, k), dl)
                assert ok, f"mode {mode} root {v0}: delta rel err {err}"
                blvl, bsig = e.batch_levels(k)
                assert np.array_equal(blvl, lvl) and np.array_equal(bsig, sig), f"mode {mode} root {v0}"
                assert stats[k] == ost, f"mode {mode} root {v0}: {stats[k]} != {ost}"
            ok, err = close(e.bc(), bc_ref)
            assert ok, f"mode {mode}: bc rel err {err}"


def test_stop_inside_batch(require_gpu):
    g = Graph.kronecker(10, 16, seed=(1, 2))
    deg = g.degrees()
    roots = np.random.default_rng(3).choice(np.flatnonzero(deg > 0), size=64, replace=False)
    probs = np.full(64, 1.0 / g.N)
    checks = np.array([int(np.argmax(deg))], np.uint32)
    bc_ref = np.zeros(g.N)
    stop_k = None
    for k, v0 in enumerate(roots):
        _, _, _, bc_ref, _ = oracle.pivot(g, int(v0), probs[k], bc_ref)
        if bc_ref[checks[0]] / 2.0 >= g.N:
            stop_k = k
            break
    assert stop_k is not None and stop_k < 63
    with Engine(g) as e:
        _mode(e, 2)
        e.set_checks(checks)
        applied, _ = run_batch(e, roots, probs, stop_c=1)
        assert applied == stop_k + 1
        ok, err = close(e.bc(), bc_ref)
        assert ok, err


def test_automatic_choice_on_a_deep_lattice(require_gpu):
    """A 200^2 lattice with 3 % of its edges deleted (~400 levels): the
    first batch runs bit-packed, the next ones as workgroup sweeps (mode 0
    switches once a batch ran more than 256 levels); utility and bc against
    the oracle."""
    g = Graph.lattice(200, 0.03, seed=3, heuristic=0)
    deg = g.degrees()
    roots = np.random.default_rng(5).choice(np.flatnonzero(deg > 0), size=3 * 64, replace=False)
    probs = np.full(64, 1.0 / g.N)
    L = _ffi.hip()
    verts = np.random.default_rng(6).choice(g.N, size=500, replace=False).astype(np.uint32)
    wts = np.random.default_rng(7).uniform(0.5, 1.5, size=500)
    bc_ref = np.zeros(g.N)
    with Engine(g) as e:
        assert L.axbc_kernel_timing(e.ctx, 1) == 0
        for b0 in range(0, len(roots), 64):
            _, stats = run_batch(e, roots[b0:b0 + 64], probs)
            if b0 == 0:
                assert max(s["nlevels"] for s in stats) > 256
            out = np.zeros(64)
            assert L.axbc_batch_utility(e.ctx, verts.ctypes.data, wts.ctypes.data, len(verts), out.ctypes.data) == 0
            for k, v0 in enumerate(roots[b0:b0 + 64]):
                _, _, dl, bc_ref, _ = oracle.pivot(g, int(v0), 1.0 / g.N, bc_ref)
                want = float(np.dot(wts, dl[verts]))
                assert abs(out[k] - want) <= 1e-9 * max(abs(want), 1.0), (b0, k)
        ms, n, b = C.c_double(), C.c_uint64(), C.c_uint64()
        assert L.axbc_kernel_stats(e.ctx, 16, C.byref(ms), C.byref(n), C.byref(b)) == 0
        assert n.value == 2, f"workgroup sweeps: {n.value}"
        ok, err = close(e.bc(), bc_ref)
        assert ok, err


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_width_256_in_every_mode(require_gpu, mode):
    """Width 256 in every mode: a workgroup sweep of all the roots (mode 2),
    or two bit-packed sweeps side by side (mode 1, and mode 0 on a graph
    of few levels); bc against the oracle either way."""
    g = Graph.lattice(32, 0.03, seed=3, heuristic=0)
    deg = g.degrees()
    roots = np.random.default_rng(9).choice(np.flatnonzero(deg > 0), size=200, replace=False)
    L = _ffi.hip()
    bc_ref = np.zeros(g.N)
    for v0 in roots:
        bc_ref = oracle.pivot(g, int(v0), 1.0 / g.N, bc_ref)[3]
    with Engine(g) as e:
        _mode(e, mode)
        assert L.axbc_set_batch_width(e.ctx, 256) == 0
        assert L.axbc_kernel_timing(e.ctx, 1) == 0
        applied, _ = run_batch(e, roots, np.full(len(roots), 1.0 / g.N))
        assert applied == len(roots)
        ms, n, b = C.c_double(), C.c_uint64(), C.c_uint64()
        assert L.axbc_kernel_stats(e.ctx, 16, C.byref(ms), C.byref(n), C.byref(b)) == 0
        assert n.value == (1 if mode == 2 else 0), f"workgroup sweeps: {n.value}"
        ok, err = close(e.bc(), bc_ref)
        assert ok, err


@pytest.mark.parametrize("stop_at", [None, 40, 170])
def test_batches_of_256(require_gpu, stop_at):
    """Width 256: one workgroup sweep of up to 256 roots, applied in two
    halves of draw order.  200 roots on a 96^2 lattice (and a Kronecker
    graph): every root's levels, path counts, delta and counters, the
    utility of all of them and the bc against the oracle; with a stop test
    that holds after root 40 (first half) or 170 (second half), exactly the
    roots up to it applied."""
    lat = Graph.lattice(96, 0.03, seed=3, heuristic=0)
    kron = Graph.kronecker(10, 16, seed=(1, 2), heuristic=1)
    for g in ((lat, kron) if stop_at is None else (lat,)):
        deg = g.degrees()
        roots = np.random.default_rng(9).choice(np.flatnonzero(deg > 0), size=200, replace=False)
        probs = np.full(200, 1.0 / g.N)
        refs = []
        bc_ref = np.zeros(g.N)
        bcs = []
        for v0 in roots:
            lvl, sig, dl, bc_ref, ost = oracle.pivot(g, int(v0), 1.0 / g.N, bc_ref)
            refs.append((lvl, sig, dl, ost))
            bcs.append(bc_ref.copy())
        checks, stop_c, want = None, 0, None
        if stop_at is not None:
            # a check vertex whose bc grows at root stop_at; probabilities
            # scaled so that bc/2 crosses c*N (c = 1) between roots stop_at-1
            # and stop_at (the lattice has no reach terms: bc scales with 1/p)
            a = next(int(v) for v in np.argsort(-bcs[-1]) if bcs[stop_at][v] > bcs[stop_at - 1][v])
            mid = 0.5 * (bcs[stop_at][a] + bcs[stop_at - 1][a])
            probs = probs * (mid / 2.0) / g.N
            bcs = []
            bc_ref = np.zeros(g.N)
            for v0 in roots:
                _, _, _, bc_ref, _ = oracle.pivot(g, int(v0), probs[0], bc_ref)
                bcs.append(bc_ref.copy())
            want = next(k for k in range(200) if bcs[k][a] / 2.0 >= g.N)
            assert want == stop_at
            checks, stop_c = np.array([a], np.uint32), 1
        with Engine(g) as e:
            _mode(e, 2)
            assert _ffi.hip().axbc_set_batch_width(e.ctx, 256) == 0
            if checks is not None:
                e.set_checks(checks)
            applied, stats = run_batch(e, roots, probs, stop_c=stop_c)
            assert applied == (200 if want is None else want + 1), (applied, want)
            for k in range(200):
                lvl, sig, dl, ost = refs[k]
                assert stats[k] == ost, f"root #{k}: {stats[k]} != {ost}"
                blvl, bsig = e.batch_levels(k)
                assert np.array_equal(blvl, lvl) and np.array_equal(bsig, sig), f"root #{k}"
                if checks is None:
                    ok, err = close(delta_of(e, k), dl)
                    assert ok, f"root #{k}: delta rel err {err}"
            if checks is None:
                ok, err = close(e.bc(), bcs[-1])
                assert ok, f"bc rel err {err}"
                verts = np.random.default_rng(3).choice(g.N, size=300, replace=False).astype(np.uint32)
                wts = np.random.default_rng(4).uniform(0.5, 1.5, size=300)
                out = np.zeros(256)
                assert _ffi.hip().axbc_batch_utility(e.ctx, verts.ctypes.data, wts.ctypes.data, 300,
                                                     out.ctypes.data) == 0
                for k in range(200):
                    want_u = float(np.dot(wts, refs[k][2][verts]))
                    assert abs(out[k] - want_u) <= 1e-9 * max(abs(want_u), 1.0), (k, out[k], want_u)
            else:
                ok, err = close(e.bc(), bcs[applied - 1])
                assert ok, f"bc after the stop: rel err {err}"


@pytest.mark.parametrize("mode", [1, 2])
def test_roots_without_edges(require_gpu, mode):
    """Roots of degree 0 in a batch (the ABI takes any vertex; the pivot
    loop skips them): a one-vertex round, delta = reach, counters as the
    oracle's, next to ordinary roots."""
    n = 400
    e = [(i, i + 1) for i in range(0, 300)] + [(0, k) for k in range(301, 340)]
    g = Graph.from_edges(n, np.array(e), heuristic=0)
    assert (g.degrees()[350:] == 0).all()
    check_batch(g, np.array([350, 0, 399, 150, 360], np.uint32), 64, f"isolated roots mode {mode}", mode=mode)
