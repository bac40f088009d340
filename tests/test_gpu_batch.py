"""Bit-packed multi-source rounds against the CPU oracle: a batch of roots
must give the same per-root delta and the same accumulated bc as the
sequence of one-root oracle rounds (delta/bc within 1e-9 relative), and the
same per-root counters."""
import ctypes as C

import numpy as np
import pytest

from axbc_amd import Engine, Graph, _ffi
from oracle import oracle
from tests.parity import close

pytestmark = pytest.mark.gpu


def run_batch(e, roots, probs, stop_c=0):
    L = _ffi.hip()
    v = np.ascontiguousarray(roots, np.uint32)
    p = np.ascontiguousarray(probs, np.float64)
    st = (_ffi.PivotStats * len(v))()
    applied = C.c_int(0)
    rc = L.axbc_batch(e.ctx, v, p, len(v), stop_c, C.byref(applied), C.cast(st, C.c_void_p))
    assert rc == 0, L.axbc_last_error()
    return applied.value, [s.as_dict() for s in st]


def delta_of(e, k):
    out = np.empty(e.N, np.float64)
    assert _ffi.hip().axbc_batch_delta(e.ctx, k, out) == 0
    return out


def check_batch(g, roots, width, label, batches=1, pull=None, chunks=None, mode=None):
    print(f"[{label}] N={g.N} ned={g.ned} width={width} roots={len(roots)} pull={pull} chunks={chunks} mode={mode}",
          flush=True)
    bc_ref = np.zeros(g.N)
    probs = np.full(len(roots), 1.0 / g.N)
    with Engine(g) as e:
        assert _ffi.hip().axbc_set_batch_width(e.ctx, width) == 0
        if mode is not None:
            assert _ffi.hip().axbc_set_batch_mode(e.ctx, mode) == 0
        if pull is not None:
            assert _ffi.hip().axbc_set_batch_pull(e.ctx, pull) == 0
        step = (len(roots) + batches - 1) // batches
        for bi, b0 in enumerate(range(0, len(roots), step)):
            if chunks is not None:
                assert _ffi.hip().axbc_set_batch_chunk(e.ctx, chunks[bi]) == 0
            rb = roots[b0:b0 + step]
            applied, stats = run_batch(e, rb, probs[b0:b0 + step])
            assert applied == len(rb)
            for k, v0 in enumerate(rb):
                lvl, sig, dl, bc_ref, ost = oracle.pivot(g, int(v0), 1.0 / g.N, bc_ref)
                ok, err = close(delta_of(e, k), dl)
                assert ok, f"{label} root #{b0 + k} ({v0}): delta rel err {err}"
                blvl, bsig = e.batch_levels(k)
                assert np.array_equal(blvl, lvl), f"{label} root #{b0 + k} ({v0}): lvl differs"
                assert np.array_equal(bsig, sig), f"{label} root #{b0 + k} ({v0}): sigma differs"
                for key in ("nvisited", "nlevels", "e_r", "e_dag", "all"):
                    assert stats[k][key] == ost[key], f"{label} root {v0}: {key} {stats[k][key]} != {ost[key]}"
            ok, err = close(e.bc(), bc_ref)
            assert ok, f"{label} after batch at {b0}: bc rel err {err}"


def test_require(require_gpu):
    pass


# pull ratio (x16): 0 = every level top-down, 32 = the default switch,
# 65536 = bottom-up from level 2 on
@pytest.mark.parametrize("pull", [0, 32, 65536])
@pytest.mark.parametrize("width", [64, 128])
@pytest.mark.parametrize("scale,seed,undirected,heur", [
    (10, (1, 2), True, 0),
    (11, (23, 24), True, 1),
    (10, (3, 4), False, 0),
])
def test_kronecker_batches(require_gpu, width, scale, seed, undirected, heur, pull):
    g = Graph.kronecker(scale, 16, seed=seed, undirected=undirected, heuristic=heur)
    deg = g.degrees()
    rng = np.random.default_rng(scale + width)
    cand = np.flatnonzero(deg > 0)
    roots = rng.choice(cand, size=min(len(cand), width + width // 2), replace=False)
    check_batch(g, roots, width, f"kron s{scale} H{heur} U{int(not undirected)}", batches=2, pull=pull)


@pytest.mark.parametrize("pull", [0, 65536])
def test_split_vertices_and_long_paths(require_gpu, pull):
    # a star centre with 5000 leaves plus a long path: split chunks in push,
    # sigma (or bottom-up) and dependency passes, and ~2000 levels
    n = 7000
    e = [(0, i) for i in range(1, 5000)] + [(i, i + 1) for i in range(4999, n - 1)]
    g = Graph.from_edges(n, np.array(e))
    roots = np.array([0, 1, 2, 4999, 6999, 5500, 3], np.uint32)
    check_batch(g, roots, 64, "star+path", pull=pull)


@pytest.mark.parametrize("width", [64, 128])
def test_bottom_up_directed_hubs(require_gpu, width):
    # directed Kronecker graph with in-hubs above the chunk size, every level
    # after the first bottom-up: split pull chunks append through NX / mark
    g = Graph.kronecker(12, 16, seed=(5, 6), undirected=False, heuristic=0)
    deg = g.degrees()
    rng = np.random.default_rng(7)
    roots = rng.choice(np.flatnonzero(deg > 0), size=width, replace=False)
    check_batch(g, roots, width, "kron s12 directed bottom-up", pull=65536)


@pytest.mark.parametrize("pull", [0, 65536])
def test_split_threshold_changes_between_batches(require_gpu, pull):
    """Split items add into path-count rows that start at zero, and a batch
    clears only the rows of the vertices the previous batch reached and that
    are above the threshold it runs at.  Two disjoint Kronecker graphs A and
    B; batches alternate between roots in A and roots in B at thresholds
    1024, 2048, 64, 128, 64: at the third batch, A's vertices of 65..1024
    edges become split while their rows still hold the first batch's counts
    (B's batch in between did not reach them), and A's delta rows must read
    reach while B's batch runs.  Top-down and bottom-up, every delta and the
    bc against the oracle."""
    from axbc_amd import kronecker_edge_list
    ea = kronecker_edge_list(11, 16 << 11, 5, 6)
    eb = kronecker_edge_list(11, 16 << 11, 7, 8) + (1 << 11)
    g = Graph.from_edges(1 << 12, np.concatenate([ea, eb]), heuristic=1)
    deg = np.diff(g.col)
    assert ((deg[: 1 << 11] > 64) & (deg[: 1 << 11] <= 1024)).sum() > 20
    rng = np.random.default_rng(11)
    ra = rng.choice(np.flatnonzero(deg[: 1 << 11] > 0), size=3 * 48, replace=False)
    rb = rng.choice(np.flatnonzero(deg[1 << 11:] > 0), size=2 * 48, replace=False) + (1 << 11)
    roots = np.concatenate([ra[:48], rb[:48], ra[48:96], rb[48:], ra[96:]])
    check_batch(g, roots, 64, "two kron s11, chunk changes", batches=5, pull=pull,
                chunks=[1024, 2048, 64, 128, 64])


@pytest.mark.parametrize("mirrored", [True, False])
def test_repeated_edges(require_gpu, mirrored):
    """A CSC handed over through the ABI may repeat an edge, and the path
    counts then count it twice (per-edge sums, as scan_col_mono2 adds them).
    Level 1 takes sigma = 1 per root only for lists without repeats, so
    these graphs must take the scan.  Every 7th vertex repeats its first
    neighbour (mirrored: the neighbour repeats it back)."""
    base = Graph.kronecker(9, 8, seed=(1, 2))
    lists = [list(base.row[base.col[v]:base.col[v + 1]]) for v in range(base.N)]
    for v in range(0, base.N, 7):
        if lists[v]:
            w = lists[v][0]
            lists[v].insert(0, w)
            if mirrored:
                lists[w].append(v)
                lists[w].sort()
    col = np.zeros(base.N + 1, np.uint32)
    col[1:] = np.cumsum([len(x) for x in lists])
    row = np.array([u for x in lists for u in x], np.uint32)
    g = Graph(base.N, col, row, np.zeros(base.N, np.uint32))
    deg = np.diff(col)
    rng = np.random.default_rng(4)
    roots = rng.choice(np.flatnonzero(deg > 0), size=96, replace=False)
    check_batch(g, roots, 64, f"repeated edges mirrored={mirrored}", batches=2)


def test_partial_batch_and_single_level_roots(require_gpu):
    n = 40  # star: the centre's round has a single level
    g = Graph.from_edges(n, np.array([(0, i) for i in range(1, n)]))
    check_batch(g, np.array([0, 5, 7], np.uint32), 64, "star")


def test_stop_inside_batch(require_gpu):
    g = Graph.kronecker(10, 16, seed=(1, 2))
    deg = g.degrees()
    rng = np.random.default_rng(3)
    roots = rng.choice(np.flatnonzero(deg > 0), size=64, replace=False)
    probs = np.full(64, 1.0 / g.N)
    checks = np.array([int(np.argmax(deg))], np.uint32)  # a hub: satisfied early
    c = 1
    # the oracle sequence: stop after the first root k with bc[a]/2 >= c*N
    bc_ref = np.zeros(g.N)
    stop_k = None
    for k, v0 in enumerate(roots):
        _, _, _, bc_ref, _ = oracle.pivot(g, int(v0), probs[k], bc_ref)
        if bc_ref[checks[0]] / 2.0 >= c * g.N:
            stop_k = k
            break
    assert stop_k is not None and stop_k < 63
    with Engine(g) as e:
        e.set_checks(checks)
        applied, _ = run_batch(e, roots, probs, stop_c=c)
        assert applied == stop_k + 1
        ok, err = close(e.bc(), bc_ref)
        assert ok, err


@pytest.mark.parametrize("width", [64, 128])
def test_graph_replays_match_host_launches(require_gpu, width):
    """The level loop replayed from captured graphs (the default) and the
    host-launched loop give bit-identical bc and delta over several batches
    (graph replays with eager work and syncs between them), on a
    high-diameter graph (a long path plus a grid) and a Kronecker graph."""
    n = 3000
    e = [(i, i + 1) for i in range(n - 1)] + [(n + r * 30 + c, n + r * 30 + c + 1) for r in range(30) for c in range(29)]
    e += [(n + r * 30 + c, n + (r + 1) * 30 + c) for r in range(29) for c in range(30)]
    graphs = [Graph.from_edges(n + 900, np.array(e)), Graph.kronecker(11, 16, seed=(9, 10))]
    for g in graphs:
        deg = g.degrees()
        rng = np.random.default_rng(width)
        roots = rng.choice(np.flatnonzero(deg > 0), size=3 * width, replace=False).astype(np.uint32)
        probs = np.full(width, 1.0 / g.N)
        out = []
        for use_graphs in (1, 0):
            with Engine(g) as eng:
                L = _ffi.hip()
                assert L.axbc_set_batch_width(eng.ctx, width) == 0
                assert L.axbc_set_batch_graphs(eng.ctx, use_graphs) == 0
                deltas = []
                for b0 in range(0, len(roots), width):
                    run_batch(eng, roots[b0:b0 + width], probs)
                    eng.sync()
                    deltas.append(delta_of(eng, width - 1))
                out.append((eng.bc(), deltas))
        print(f"[graphs] N={g.N} width={width}", flush=True)
        np.testing.assert_array_equal(out[0][0], out[1][0])
        for a, b in zip(out[0][1], out[1][1]):
            np.testing.assert_array_equal(a, b)
        # and against the oracle
        bc_ref = np.zeros(g.N)
        for v0 in roots:
            _, _, _, bc_ref, _ = oracle.pivot(g, int(v0), 1.0 / g.N, bc_ref)
        ok, err = close(out[0][0], bc_ref)
        assert ok, err


def test_utility_per_root_in_draw_order(require_gpu):
    # roots of very different degrees sit in lanes by degree; the utility of
    # root k must still be sum_v weight(v) * delta_k(v) of the k-th root drawn
    g = Graph.kronecker(11, 16, seed=(3, 4), heuristic=0)
    deg = g.degrees()
    order = np.argsort(deg, kind="stable")
    cand = order[deg[order] > 0]
    roots = np.concatenate([cand[:20], cand[-20:], cand[len(cand) // 2:len(cand) // 2 + 24]])
    rng = np.random.default_rng(7)
    rng.shuffle(roots)
    verts = rng.choice(g.N, size=300, replace=False).astype(np.uint32)
    wts = rng.uniform(0.1, 2.0, size=300)
    with Engine(g) as e:
        assert _ffi.hip().axbc_set_batch_width(e.ctx, 64) == 0
        applied, _ = run_batch(e, roots, np.full(len(roots), 1.0 / g.N))
        assert applied == len(roots)
        out = np.zeros(64, np.float64)
        assert _ffi.hip().axbc_batch_utility(e.ctx, verts.ctypes.data, wts.ctypes.data, len(verts),
                                             out.ctypes.data) == 0
        bc_ref = np.zeros(g.N)
        for k, v0 in enumerate(roots):
            _, _, dl, bc_ref, _ = oracle.pivot(g, int(v0), 1.0 / g.N, bc_ref)
            want = float(np.dot(wts, dl[verts]))
            assert abs(out[k] - want) <= 1e-9 * max(abs(want), 1.0), (k, out[k], want)


@pytest.mark.parametrize("width", [64, 128])
def test_sigma_wraps_to_zero_batched(require_gpu, width):
    # sigma = 2^32 = 0 mod 2^32 on layer 4: inf/NaN dependencies must come
    # out at the same positions as the one-root oracle rounds
    from tests.test_gpu_plain import layered
    g = layered(256, 6)
    roots = np.array([0, 1 + 5 * 256 + 7, 1 + 2 * 256 + 3, 1 + 4 * 256, 5], np.uint32)
    check_batch(g, roots, width, "sigma wraps to 0")


def _persist(e, mode, limit=10.0):
    assert _ffi.hip().axbc_set_batch_persistent(e.ctx, mode, limit) == 0, _ffi.hip().axbc_last_error()


@pytest.mark.parametrize("width", [64, 128])
def test_persistent_level_loop(require_gpu, width):
    """High-diameter graphs (a 96^2 lattice with 3% of its edges deleted, a
    long path with pendants): the sweep in persistent cooperative launches
    (mode 2) against the oracle, and bit-identical to the launch-per-level
    loop (mode 0)."""
    lat = Graph.lattice(96, 0.03, seed=3, heuristic=0)
    n = 1500
    path = Graph.from_edges(n, np.array([(i, i + 1) for i in range(n - 1)] + [(i, n - 1 - i // 2) for i in range(0, 40, 3)]))
    for g, label in ((lat, "lattice 96"), (path, "path 1500")):
        deg = g.degrees()
        rng = np.random.default_rng(width)
        roots = rng.choice(np.flatnonzero(deg > 0), size=width - 5, replace=False)
        probs = np.full(len(roots), 1.0 / g.N)
        out = {}
        for mode in (2, 0):
            bc_ref = np.zeros(g.N)
            with Engine(g) as e:
                assert _ffi.hip().axbc_set_batch_width(e.ctx, width) == 0
                _persist(e, mode)
                applied, stats = run_batch(e, roots, probs)
                assert applied == len(roots)
                if mode == 2:
                    for k in (0, len(roots) // 2, len(roots) - 1):
                        lvl, sig, dl, bc_ref, ost = oracle.pivot(g, int(roots[k]), 1.0 / g.N, np.zeros(g.N))
                        blvl, bsig = e.batch_levels(k)
                        assert np.array_equal(blvl, lvl) and np.array_equal(bsig, sig), f"{label} root #{k}"
                        ok, err = close(delta_of(e, k), dl)
                        assert ok, f"{label} root #{k}: delta rel err {err}"
                        assert stats[k] == ost, f"{label} root #{k}: {stats[k]} != {ost}"
                    assert max(s["nlevels"] for s in stats) > 60
                out[mode] = e.bc()
        assert np.array_equal(out[2], out[0]), f"{label}: persistent and per-level bc differ"


def test_persistent_barrier_limit_fails_cleanly(require_gpu):
    """A barrier limit far below the time a level takes (forced from the
    host): the launch ends in every workgroup, the batch reports the error,
    bc is unchanged, and the next batch with the normal limit rebuilds its
    state and matches the oracle."""
    g = Graph.lattice(64, 0.03, seed=3, heuristic=0)
    deg = g.degrees()
    roots = np.random.default_rng(1).choice(np.flatnonzero(deg > 0), size=40, replace=False)
    probs = np.full(len(roots), 1.0 / g.N)
    L = _ffi.hip()
    with Engine(g) as e:
        _persist(e, 2, 1e-9)
        rc = L.axbc_batch(e.ctx, np.ascontiguousarray(roots, np.uint32), probs, len(roots), 0, C.byref(C.c_int(0)), None)
        assert rc == -3 and b"limit" in L.axbc_last_error()
        assert (e.bc() == 0).all()
        _persist(e, 2, 10.0)
        applied, _ = run_batch(e, roots, probs)
        assert applied == len(roots)
        bc_ref = np.zeros(g.N)
        for v in roots:
            _, _, _, bc_ref, _ = oracle.pivot(g, int(v), 1.0 / g.N, bc_ref)
        ok, err = close(e.bc(), bc_ref)
        assert ok, err
