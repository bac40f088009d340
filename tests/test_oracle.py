"""Pins of the CPU oracle (oracle/brandes_oracle.c) by brute force, closed
forms and invariants, so that the oracle can check the HIP path."""
import math

import numpy as np
import pytest

from axbc_amd import Graph
from oracle import oracle
from tests import bruteforce as bf


def rand_graph(rng, n, m, undirected=True):
    e = rng.integers(0, n, size=(m, 2))
    return Graph.from_edges(n, e, undirected=undirected)


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("undirected", [True, False])
def test_bruteforce_small(seed, undirected):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 12))
    g = rand_graph(rng, n, int(rng.integers(n, 3 * n)), undirected)
    adj = bf.adjacency(g.N, g.col, g.row)
    bc_expect = np.zeros(n)
    bc = np.zeros(n)
    for s in range(n):
        if g.col[s + 1] == g.col[s]:
            continue
        p = 1.0 / n
        lvl, sig, dl, bc, st = oracle.pivot(g, s, p, bc)
        dist, cnt = bf.bfs(adj, s)
        assert list(lvl) == dist
        assert [int(x) for x in sig] == [c % (1 << 32) for c in cnt]
        dep = bf.dependencies(adj, s)
        nlev = max(dist)
        for v in range(n):
            if v == s:
                # the root's dependency is only accumulated when the BFS
                # has a single level (bc_func_mono's do-while); then every
                # level-1 vertex contributes (0 + 1) / 1 * 1
                exp = float(sum(1 for t in range(n) if dist[t] == 1)) if nlev == 1 else 0.0
                assert dl[v] == exp, (s, v)
            else:
                assert dl[v] == pytest.approx(dep[v], rel=1e-12, abs=1e-12), (s, v)
                bc_expect[v] += dep[v] / p
        assert st["nvisited"] == sum(1 for d in dist if d >= 0)
        assert st["nlevels"] == nlev
    np.testing.assert_allclose(bc, bc_expect, rtol=1e-12, atol=1e-9)


def path(n):
    return Graph.from_edges(n, np.array([(i, i + 1) for i in range(n - 1)]))


def test_path_closed_form():
    n = 50
    g = path(n)
    lvl, sig, dl, _, st = oracle.pivot(g, 0, 1.0)
    assert list(lvl) == list(range(n))
    assert (sig == 1).all()
    assert list(dl[1:]) == [float(n - 1 - i) for i in range(1, n)]
    assert dl[0] == 0.0  # more than one level: the root keeps reach = 0
    assert st["e_dag"] == n - 1 and st["e_r"] == 2 * (n - 1)


def test_star_closed_form():
    n = 40
    g = Graph.from_edges(n, np.array([(0, i) for i in range(1, n)]))
    # from the centre: one level, the root's delta is the do-while case
    lvl, sig, dl, _, st = oracle.pivot(g, 0, 1.0)
    assert st["nlevels"] == 1 and dl[0] == n - 1 and (dl[1:] == 0).all()
    # from a leaf: the centre lies on the paths to the other n-2 leaves
    lvl, sig, dl, _, _ = oracle.pivot(g, 5, 1.0)
    assert dl[0] == n - 2


def test_cycle_antipode_two_paths():
    k = 9
    n = 2 * k
    g = Graph.from_edges(n, np.array([(i, (i + 1) % n) for i in range(n)]))
    lvl, sig, dl, _, _ = oracle.pivot(g, 0, 1.0)
    assert lvl[k] == k and sig[k] == 2
    assert (sig >= 1).all()
    # every non-antipodal vertex at distance j carries the half of the
    # antipode's paths plus the k-1-j vertices behind it
    for j in range(1, k):
        assert dl[j] == pytest.approx((k - 1 - j) + 0.5)


def test_complete_graph():
    n = 12
    e = [(i, j) for i in range(n) for j in range(i + 1, n)]
    g = Graph.from_edges(n, np.array(e))
    lvl, sig, dl, _, st = oracle.pivot(g, 3, 1.0)
    assert (sig == 1).all() and st["nlevels"] == 1
    assert (dl[np.arange(n) != 3] == 0).all()


def test_ladder_binomial_counts():
    k = 20  # 2 x k grid: paths from a corner to (r, c) = binom(c + r, r)
    idx = lambda r, c: r * k + c  # noqa: E731
    e = [(idx(r, c), idx(r, c + 1)) for r in range(2) for c in range(k - 1)]
    e += [(idx(0, c), idx(1, c)) for c in range(k)]
    g = Graph.from_edges(2 * k, np.array(e))
    lvl, sig, _, _, _ = oracle.pivot(g, idx(0, 0), 1.0)
    for c in range(k):
        assert sig[idx(0, c)] == 1
        assert sig[idx(1, c)] == c + 1


def test_sigma_wraps_mod_2_32():
    # layers of a vertices, consecutive layers completely connected: sigma of
    # layer j is a^j, which passes 2^32 and must wrap like the uint32 atomics
    a, layers = 3, 24
    e = []
    for j in range(layers - 1):
        for x in range(a):
            for y in range(a):
                e.append((1 + j * a + x, 1 + (j + 1) * a + y))
    e += [(0, 1 + x) for x in range(a)]
    n = 1 + layers * a
    g = Graph.from_edges(n, np.array(e))
    lvl, sig, dl, _, _ = oracle.pivot(g, 0, 1.0)
    for j in range(layers):
        for x in range(a):
            assert int(sig[1 + j * a + x]) == pow(a, j, 1 << 32)
    assert a ** (layers - 1) > (1 << 32)


@pytest.mark.parametrize("scale,seed", [(9, (1, 2)), (10, (23, 24))])
def test_dependency_sum_invariant(scale, seed):
    # sum_v delta_s(v) = sum_t (d(s,t) - 1) over reached t != s
    g = Graph.kronecker(scale, 16, seed=seed)
    deg = g.degrees()
    for s in np.flatnonzero(deg > 0)[:6]:
        lvl, sig, dl, _, st = oracle.pivot(g, int(s), 1.0)
        reached = lvl > 0
        expect = float((lvl[reached] - 1).sum())
        got = float(dl[np.arange(g.N) != s].sum())
        assert got == pytest.approx(expect, rel=1e-9)


def test_lim_x_sample_wraps():
    L = oracle.lib()
    assert L.oracle_lim_x_sample(1 << 12) == 1 << 24
    assert L.oracle_lim_x_sample(1 << 16) == 0
    assert L.oracle_lim_x_sample((1 << 24) - 13) == ((((1 << 24) - 13) ** 2) % (1 << 32))


def test_fixed_point_weights_exact_and_monotone():
    rng = np.random.default_rng(5)
    w = rng.random(1000) * 1e3
    w[::7] = 0.0
    W = np.zeros(1000, np.uint64)
    total = oracle.lib().oracle_fixed_weights(1000, w, W)
    assert total == int(W.astype(object).sum())
    assert total < (1 << 63)
    # the shift is one power of two for the whole array
    scale = 2.0 ** round(math.log2(float(W.max()) / w.max()))
    frac = w * scale - W.astype(np.float64)
    assert ((frac > -1e-3) & (frac < 1.0 + 1e-3)).all()  # W = floor(w * 2^s)


def test_weighted_sample_first_reaching_prefix():
    W = np.array([0, 5, 0, 10, 5], np.uint64)
    p = oracle.C.c_double()
    L = oracle.lib()
    total = 20
    # draw 0 -> r = 0 -> first prefix >= 0 is index 0 (weight 0)
    assert L.oracle_weighted_sample(5, W, total, 0, oracle.C.byref(p)) == 0
    # draw just below half -> r ~ 10 -> prefix [0,5,5,15,20] reaches at 3
    d = int(2147483647 * 0.5) - 1
    assert L.oracle_weighted_sample(5, W, total, d, oracle.C.byref(p)) == 3
    assert p.value == 10 / 20


def test_hdn_sample_probability():
    # path 0-1-2: deg = 1,2,1; deg^2 prefix = 1,5,6
    g = path(3)
    p = oracle.C.c_double()
    v = oracle.lib().oracle_hdn_sample(3, g.col, g.row, 0, oracle.C.byref(p))
    # r = 0 -> u = 0 (degree 1) -> its only neighbour 1
    assert v == 1 and p.value == pytest.approx(2 / 6)


def test_run_uniform_draw_order():
    # Pivot loop with c = 0 (no stop test): x check draws, one discarded
    # initial draw, then one lrand48() % (N+1) per attempt; a repeated root
    # is redrawn inside the same iteration (LIMxSAMPLE = 2^20 at N = 2^10),
    # a degree-0 root consumes the iteration.  The draws come from libc's
    # lrand48, itself pinned against the POSIX recurrence in
    # test_lrand48_posix_recurrence.
    import ctypes
    g = Graph.kronecker(10, 16, seed=(1, 2))
    nbfs, x = 60, 3
    pv, pr, bc, rs = oracle.run(g, nbfs, c=0, x=x, z=0)
    libc = ctypes.CDLL(None)
    libc.lrand48.restype = ctypes.c_long
    libc.srand48(ctypes.c_long(4343543))
    N = g.N
    for _ in range(x + 1):
        libc.lrand48()
    deg = g.degrees()
    seen, expect, skipped = set(), [], 0
    for _ in range(nbfs):
        v = libc.lrand48() % (N + 1)
        while v < N and v in seen:
            v = libc.lrand48() % (N + 1)
        if v >= N or deg[v] == 0:
            skipped += 1
            if v < N:
                seen.add(v)
            continue
        seen.add(v)
        expect.append(v)
    assert list(pv) == expect
    assert rs["skipped"] == skipped and rs["nrounds"] == len(expect)
    assert (pr == 1.0 / N).all()


def test_lrand48_posix_recurrence():
    # X_{n+1} = (0x5DEECE66D X_n + 0xB) mod 2^48, srand48: X = seed << 16 |
    # 0x330E, lrand48 = X >> 17
    import ctypes
    libc = ctypes.CDLL(None)
    libc.lrand48.restype = ctypes.c_long
    libc.srand48(ctypes.c_long(4343543))
    X = (4343543 << 16) | 0x330E
    for _ in range(1000):
        X = (0x5DEECE66D * X + 0xB) % (1 << 48)
        assert libc.lrand48() == X >> 17


def test_oracle_generator_is_lrand48():
    # the oracle's per-thread generator draws libc's srand48/lrand48 sequence
    import ctypes
    libc = ctypes.CDLL(None)
    libc.lrand48.restype = ctypes.c_long
    L = oracle.lib()
    L.oracle_lrand48.restype = ctypes.c_long
    for seed in (0, 1, 4343543, 2**31 - 1, -5):
        libc.srand48(ctypes.c_long(seed))
        L.oracle_srand48(ctypes.c_long(seed))
        for _ in range(200):
            assert L.oracle_lrand48() == libc.lrand48()


def test_stop_test_stops():
    g = Graph.kronecker(8, 8, seed=(1, 2))
    pv, pr, bc, rs = oracle.run(g, g.N, c=1, x=1, z=0)
    # with c = 1 and one check vertex the loop may stop early; whenever it
    # stopped the criterion holds
    if rs["stopped"]:
        assert math.isfinite(bc.sum())


def test_lcc_closed_forms():
    # triangle with a pendant: lcc(0)=lcc(1)=1 ... vertex 2 (deg 3): 2 of 6
    g = Graph.from_edges(4, np.array([(0, 1), (1, 2), (0, 2), (2, 3)]))
    lcc = np.zeros(4, np.float32)
    oracle.lib().oracle_lcc.argtypes = [oracle.C.c_uint32, oracle.u32p, oracle.u32p,
                                        np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")]
    oracle.lib().oracle_lcc(4, g.col, g.row, lcc)
    assert list(lcc) == [1.0, 1.0, np.float32(2) / np.float32(6), 0.0]
    # K_5: every vertex has lcc 1; a star centre has 0
    k5 = Graph.from_edges(5, np.array([(i, j) for i in range(5) for j in range(i + 1, 5)]))
    l5 = np.zeros(5, np.float32)
    oracle.lib().oracle_lcc(5, k5.col, k5.row, l5)
    assert (l5 == 1.0).all()


def test_lcc_matches_triangle_count():
    # lcc(v) * d(d-1) = 2 * triangles through v on a random undirected graph
    rng = np.random.default_rng(2)
    n = 60
    g = Graph.from_edges(n, rng.integers(0, n, size=(400, 2)))
    adj = [set(g.row[g.col[v]:g.col[v + 1]]) - {v} for v in range(n)]
    lcc = np.zeros(n, np.float32)
    oracle.lib().oracle_lcc(n, g.col, g.row, lcc)
    for v in range(n):
        d = g.col[v + 1] - g.col[v]
        tri = sum(1 for a in adj[v] for b in adj[v] if a < b and b in adj[a])
        exp = np.float32(0) if d < 2 else np.float32(2 * tri) / np.float32(d * (d - 1))
        assert lcc[v] == exp


@pytest.mark.parametrize("z", [2, 6])
def test_run_lcc_and_mixture(z):
    g = Graph.kronecker(9, 16, seed=(1, 2))
    pv, pr, bc, rs = oracle.run(g, 60, c=0, x=2, z=z)
    assert rs["nrounds"] > 10
    assert np.isfinite(pr).all() and (pr > 0).all()


def test_sync_batch_freezes_dynamic_distribution():
    g = Graph.kronecker(9, 16, seed=(1, 2))
    pv0, _, _, rs0 = oracle.run(g, 80, c=0, x=1, z=4)
    pv1, _, _, rs1 = oracle.run(g, 80, c=0, x=1, z=4, sync_batch=1)
    # groups of one round are the per-round rule
    assert list(pv0) == list(pv1)
    pv8, pr8, _, _ = oracle.run(g, 80, c=0, x=1, z=4, sync_batch=8)
    assert (pr8[:8] == 1.0 / g.N).all()  # first group: uniform
    assert list(pv8[:8]) == list(pv0[:1]) + list(pv8[1:8])


@pytest.mark.parametrize("z", [3, 5, 6])
def test_sync_batch_of_one_is_the_per_round_rule(z):
    # a group of one round freezes the state of that round: the exact loop,
    # pivots and probabilities bit for bit (the mixture's bc/delta terms too)
    g = Graph.kronecker(9, 16, seed=(3, 4), heuristic=1)
    pv0, pr0, bc0, rs0 = oracle.run(g, 70, c=0, x=2, z=z)
    pv1, pr1, bc1, rs1 = oracle.run(g, 70, c=0, x=2, z=z, sync_batch=1)
    assert list(pv0) == list(pv1) and rs0 == rs1
    np.testing.assert_array_equal(pr0, pr1)
    np.testing.assert_array_equal(bc0, bc1)


def test_sync_batch_mixture_freezes_bc_and_delta_terms():
    # -z 6 in groups of 16: the first group uses the round-0 rule for the
    # bc/delta components (1/N each); later probabilities differ from the
    # per-round rule only through those frozen terms
    g = Graph.kronecker(9, 16, seed=(3, 4), heuristic=1)
    pv16, pr16, _, rs16 = oracle.run(g, 70, c=0, x=2, z=6, sync_batch=16)
    pv0, pr0, _, _ = oracle.run(g, 70, c=0, x=2, z=6)
    assert rs16["nrounds"] > 32
    # component draws do not depend on bc, so the pivot sequence is the same
    assert list(pv16) == list(pv0)
    assert pr16[0] == pr0[0]
    assert np.isfinite(pr16).all() and (pr16 > 0).all()
    assert not np.array_equal(pr16[16:], pr0[16:])


def test_mlp_oracle_gradient_by_finite_differences():
    from oracle import mlp_oracle as mo
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mo.N_PARAMS)
    x = rng.normal(size=(40, 4))
    y = rng.normal(size=40)
    loss, g = mo.loss_and_grad(p, x, y, bf16_products=False)
    for i in rng.choice(mo.N_PARAMS, 60, replace=False):
        h = 1e-6
        q = p.copy()
        q[i] += h
        lp, _ = mo.loss_and_grad(q, x, y, bf16_products=False)
        q[i] -= 2 * h
        lm, _ = mo.loss_and_grad(q, x, y, bf16_products=False)
        assert (lp - lm) / (2 * h) == pytest.approx(g[i], rel=1e-4, abs=1e-7)


def test_mlp_oracle_bf16_products():
    """The matrix-core rounding (h1 into layer 2, dz2 into the backward
    products) moves the loss and gradient by about the BF16 precision and
    no more, and touches exactly the terms it should: with activations and
    gradients already on the BF16 grid it changes nothing."""
    from oracle import mlp_oracle as mo
    rng = np.random.default_rng(1)
    p = mo.to_bf16(rng.normal(0, 0.3, mo.N_PARAMS).astype(np.float32)).astype(np.float64)
    x = rng.normal(size=(64, 4))
    y = rng.normal(size=64)
    l0, g0 = mo.loss_and_grad(p, x, y, bf16_products=False)
    l1, g1 = mo.loss_and_grad(p, x, y)
    assert l1 != l0 and abs(l1 - l0) <= 2 ** -7 * abs(l0)
    assert np.abs(g1 - g0).max() <= 2 ** -6 * np.abs(g0).max()
    # the output-layer terms (b3, w3) see only the forward rounding; b2 the
    # unrounded dz2: recompute them by hand from the rounded forward pass
    s, (z1, h1, z2, h2) = mo.forward(p, x)
    ds = 2.0 * (s - y) / len(y)
    assert g1[mo.O_B3] == pytest.approx(ds.sum(), rel=1e-12)
    np.testing.assert_allclose(g1[mo.O_W3:mo.O_B3], h2.T @ ds, rtol=1e-12)
    w3 = p[mo.O_W3:mo.O_B3]
    np.testing.assert_allclose(g1[mo.O_B2:mo.O_W3], (np.outer(ds, w3) * (z2 > 0)).sum(0), rtol=1e-12)
    # one hidden unit, one sample: a network whose h1 is exactly on the BF16
    # grid (W1 = 0, b1 on the grid) gives the same forward either way
    q = p.copy()
    q[mo.O_W1:mo.O_B1] = 0.0
    s0, _ = mo.forward(q, x, bf16_products=False)
    s1, _ = mo.forward(q, x)
    np.testing.assert_array_equal(s0, s1)


def test_bf16_rounding():
    from oracle import mlp_oracle as mo
    a = np.array([1.0, 1.00390625, 1.0 + 2 ** -8 + 2 ** -10, -3.14159], np.float32)
    b = mo.to_bf16(a)
    assert b[0] == 1.0 and b[1] == 1.0  # tie to even
    assert b[2] == np.float32(1.0 + 2 ** -7)
    assert abs(b[3] + 3.140625) < 1e-6


def test_parallel_rounds_equal_sequential():
    # the OpenMP throughput form (CPU baseline) sums the same rounds
    g = Graph.kronecker(10, 16, seed=(1, 2), heuristic=1)
    deg = np.diff(g.col)
    roots = np.flatnonzero(deg > 0)[:40].astype(np.uint32)
    probs = np.full(len(roots), 1.0 / g.N)
    bc, er = np.zeros(g.N), 0
    for v, p in zip(roots, probs):
        _, _, _, bc, st = oracle.pivot(g, int(v), float(p), bc)
        er += st["e_r"]
    bp, erp = oracle.pivots_parallel(g, roots, probs, 4)
    assert erp == er
    assert np.max(np.abs(bp - bc) / np.maximum(np.abs(bc), 1e-300)) < 1e-12


def test_sigma_wrapping_to_zero_gives_infinite_dependencies():
    # layers of 256: sigma of layers 4 and 5 is 2^32, 2^40 = 0 (uint32); the
    # last layer's x = (0+1)/0 is infinite, layer 4 adds inf * sigma = inf * 0
    # = NaN, and NaN reaches every layer above
    a, layers = 256, 6
    e = [(0, 1 + x) for x in range(a)]
    for j in range(layers - 1):
        for x in range(a):
            for y in range(a):
                e.append((1 + j * a + x, 1 + (j + 1) * a + y))
    g = Graph.from_edges(1 + layers * a, np.array(e))
    lvl, sig, dl, _, _ = oracle.pivot(g, 0, 1.0)
    assert (sig[1 + 4 * a:] == 0).all()
    assert (dl[1 + 5 * a:] == 0).all()
    assert np.isnan(dl[1:1 + 5 * a]).all()


def test_loops_on_several_threads_draw_their_own_sequences():
    """The oracle's generator state is per thread: loops run at once on
    several threads give each the sequence of a loop run alone."""
    from concurrent.futures import ThreadPoolExecutor
    gs = [Graph.kronecker(9 + k % 2, 8, seed=(k + 1, k + 2), heuristic=1) for k in range(4)]
    alone = [oracle.run(g, 120, c=0, x=1, z=6, seed=100 + k) for k, g in enumerate(gs)]
    with ThreadPoolExecutor(max_workers=4) as ex:
        together = list(ex.map(lambda kg: oracle.run(kg[1], 120, c=0, x=1, z=6, seed=100 + kg[0]),
                               list(enumerate(gs)) * 2))
    for k, (pv, pr, bc, rs) in enumerate(together):
        apv, apr, abc, ars = alone[k % 4]
        assert np.array_equal(pv, apv) and np.array_equal(pr, apr) and rs == ars
        assert np.array_equal(bc, abc)
