"""Independent pins of three oracle functions that the closed forms of
test_oracle.py leave open:

* the -z 6 mixture probability and draw order (src/axbc2d.c:2798,
  2876-3051), on an 8-vertex graph whose degree, neighbour-degree and LCC
  tables are written out by hand, with bc and delta from the brute-force
  pair definition (tests/bruteforce.py) and the draws from a pure-Python
  POSIX lrand48;
* the adaptive stop rule (src/axbc2d.c:2846-2870) on a star whose centre
  is the check vertex, where the stopping round is known in closed form;
* the high-degree-neighbourhood draw over a WRAPPING uint32 prefix of
  (LOCINT)pow(deg, 2) (src/axbc2d.c:2681-2685, 2188-2198, 2895-2928), on a
  graph with vertices of degree >= 65,536, against Python integers mod 2^32
  and a linear first match.
"""
import ctypes
import math

import numpy as np
import pytest

from axbc_amd import Graph
from oracle import oracle
from tests import bruteforce as bf
from tests.lrand48 import Lrand48

SEED = 4343543
M32 = 0xFFFFFFFF


def fixed_point(w):
    """The fixed-point weights of the dynamic distributions: s = 61 - e
    with max(w) * N = m * 2^e, m in [0.5, 1); W = floor(w * 2^s)."""
    w = [float(x) for x in w]
    wmax = max([x for x in w if math.isfinite(x) and x > 0], default=0.0)
    if wmax <= 0:
        return [0] * len(w), 0
    _, e = math.frexp(wmax * len(w))
    s = 61 - e
    W = [int(math.ldexp(x, s)) if (math.isfinite(x) and x > 0) else 0 for x in w]
    return W, sum(W)


def weighted_pick(W, total, draw):
    rd = draw / 2147483647.0 * float(total)
    r = total if rd >= float(total) else int(rd)
    run = 0
    for i, x in enumerate(W):
        run += x
        if run >= r:
            return i
    return len(W) - 1


# --------------------------------------------------------------------------
# -z 6 mixture
# --------------------------------------------------------------------------

# two triangles joined by the edge 2-3, and a tail 5-6-7
MIX_EDGES = [(0, 1), (0, 2), (1, 2), (2, 3), (3, 4), (3, 5), (4, 5), (5, 6), (6, 7)]
MIX_N = 8
MIX_DEG = [2, 2, 3, 3, 2, 3, 2, 1]
MIX_SUMDEG2 = 44                           # sum of deg^2
MIX_SDN = [5, 5, 7, 8, 6, 7, 4, 2]         # sum of the neighbours' degrees
THIRD = np.float32(2) / np.float32(6)      # 2 triangles-pairs of 6 ordered pairs, in float
MIX_LCC = [np.float32(1), np.float32(1), THIRD, THIRD, np.float32(1), THIRD, np.float32(0), np.float32(0)]


def test_mixture_tables_by_hand():
    g = Graph.from_edges(MIX_N, np.array(MIX_EDGES))
    adj = bf.adjacency(g.N, g.col, g.row)
    assert [len(a) for a in adj] == MIX_DEG
    assert sum(d * d for d in MIX_DEG) == MIX_SUMDEG2 == sum(MIX_SDN)
    assert [sum(MIX_DEG[w] for w in adj[v]) for v in range(MIX_N)] == MIX_SDN
    lcc = np.zeros(MIX_N, np.float32)
    L = oracle.lib()
    L.oracle_lcc.argtypes = [ctypes.c_uint32, oracle.u32p, oracle.u32p,
                             np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")]
    L.oracle_lcc(MIX_N, g.col, g.row, lcc)
    assert list(lcc) == MIX_LCC


def simulate_mixture(nbfs, x):
    """The -z 6 pivot loop with c = 0 on the hand tables: returns (pivots,
    probabilities, skipped)."""
    g = Graph.from_edges(MIX_N, np.array(MIX_EDGES))
    adj = bf.adjacency(g.N, g.col, g.row)
    N = MIX_N
    rng = Lrand48(SEED)
    for _ in range(x):
        rng.uniform_int(N)                       # check vertices (unused at c = 0)
    mix = rng.uniform_int(4)                     # initial component, 2798
    prefix, run = [], 0
    for d in MIX_DEG:
        run = (run + d * d) & M32
        prefix.append(run)
    S = prefix[-1]
    Wl, Tl = fixed_point(MIX_LCC)
    bc = [0.0] * N
    last_delta = [0.0] * N
    seen, pivots, probs, skipped = set(), [], [], 0
    lim = (N * N) & M32
    for _ in range(min(nbfs, N)):
        bsm, good = 0, False
        v0, p = 0, 0.0
        while not good:
            if bsm > lim:
                break
            first = len(pivots) == 0
            k = mix
            if k == 0 or (k in (3, 4, 5) and first):
                v0 = rng.uniform_int(N)
            elif k == 1:
                r = rng.uniform_int(S)
                u = next((i for i in range(N) if prefix[i] >= r), N - 1)
                d = MIX_DEG[u]
                if d > 0:
                    v0 = adj[u][0 if d == 1 else rng.uniform_int(d - 1)]
                else:
                    v0 = u
            elif k == 2:
                v0 = weighted_pick(Wl, Tl, rng.next())
            else:
                raise AssertionError("components 3-5 only occur in the first round")
            inside = v0 < N
            p = 1.0 / N
            if first:
                p = p + 1.0 / N
                p = p + 1.0 / N
            else:
                Wb, Tb = fixed_point(bc)
                Wd, Td = fixed_point(last_delta)
                p = p + (Wb[v0] if inside else 0) / Tb
                p = p + (Wd[v0] if inside else 0) / Td
            p = p + (MIX_SDN[v0] if inside else 0) / S
            p = p + (Wl[v0] if inside else 0) / Tl
            p = p / 5.0
            mix = rng.uniform_int(2)
            if inside and v0 in seen:
                bsm += 1
            else:
                good = True
        if bsm > lim:
            continue
        if v0 >= N or MIX_DEG[v0] == 0:
            skipped += 1
            if v0 < N:
                seen.add(v0)
            continue
        dep = bf.dependencies(adj, v0)
        for v in range(N):
            if v != v0:
                bc[v] += dep[v] / p
        last_delta = [0.0 if v == v0 else dep[v] for v in range(N)]
        seen.add(v0)
        pivots.append(v0)
        probs.append(p)
    return pivots, probs, skipped


@pytest.mark.parametrize("x", [1, 2, 5])
def test_mixture_loop_by_hand(x):
    g = Graph.from_edges(MIX_N, np.array(MIX_EDGES))
    exp_pv, exp_pr, exp_skip = simulate_mixture(MIX_N, x)
    pv, pr, _, rs = oracle.run(g, MIX_N, c=0, x=x, z=6)
    assert list(pv) == exp_pv
    assert rs["skipped"] == exp_skip
    assert len(exp_pv) >= 5
    # later rounds use the fixed-point bc and delta weights, whose last bit
    # can differ between the brute-force and the recursive dependencies
    np.testing.assert_allclose(pr, exp_pr, rtol=1e-12, atol=0)
    assert pr[0] == exp_pr[0]


def test_mixture_first_round_terms():
    # round 0: three uniform terms (components 0, 3, 4), then HDN and LCC
    g = Graph.from_edges(MIX_N, np.array(MIX_EDGES))
    pv, pr, _, _ = oracle.run(g, 1, c=0, x=1, z=6)
    v = int(pv[0])
    Wl, Tl = fixed_point(MIX_LCC)
    want = (3.0 / MIX_N + MIX_SDN[v] / MIX_SUMDEG2 + Wl[v] / Tl) / 5.0
    assert pr[0] == pytest.approx(want, rel=1e-15)


# --------------------------------------------------------------------------
# adaptive stop rule
# --------------------------------------------------------------------------

def star_around(centre, N):
    return Graph.from_edges(N, np.array([(centre, v) for v in range(N) if v != centre]))


@pytest.mark.parametrize("c,leaf_rounds", [(21, 3), (22, 4), (28, 4), (29, 5)])
def test_stop_rule_star_closed_form(c, leaf_rounds):
    """N = 16, the centre is the one check vertex (the first lrand48 draw
    % 17 is 8, so the star is built around 8).  A leaf root gives the
    centre delta = N - 2 and bc += (N - 2) / (1/N) = 224 (exact: 1/16 is a
    power of two); a centre root adds nothing.  The loop stops at the first
    attempt whose bc[centre] / 2 >= c * N, i.e. after
    ceil(2c / (N - 2)) leaf rounds; equality stops it (c = 21, 28)."""
    N = 16
    rng = Lrand48(SEED)
    centre = rng.uniform_int(N)
    assert centre == 8
    assert leaf_rounds == math.ceil(2 * c / (N - 2))
    g = star_around(centre, N)
    pv, pr, bc, rs = oracle.run(g, N, c=c, x=1, z=0)
    assert rs["stopped"]
    assert bc[centre] == 224.0 * leaf_rounds
    assert (np.delete(bc, centre) == 0).all()
    leaves = [v for v in pv if v != centre]
    assert len(leaves) == leaf_rounds and pv[-1] != centre
    # the stop is tested before every attempt: the attempts are exactly
    # those of the rounds and skips before the stopping one
    assert rs["attempts"] == rs["nrounds"] + rs["skipped"]
    # one attempt fewer and the loop has not stopped yet
    pv2, _, bc2, rs2 = oracle.run(g, rs["attempts"], c=c, x=1, z=0)
    assert not rs2["stopped"] and list(pv2) == list(pv)


def test_stop_rule_leaf_check_never_stops():
    # check vertices 8 (the centre) and 4 (a leaf, bc stays 0): no stop
    N = 16
    g = star_around(8, N)
    pv, pr, bc, rs = oracle.run(g, N, c=1, x=2, z=0)
    assert not rs["stopped"]
    assert rs["nrounds"] + rs["skipped"] == N


def test_stop_rule_follows_python_loop():
    """The same star, c = 21: the loop by the reference's order (x check
    draws, one discarded draw, then per attempt the stop test and
    lrand48 % (N + 1) with redraws of sampled vertices) in pure Python."""
    N, c = 16, 21
    g = star_around(8, N)
    rng = Lrand48(SEED)
    checks = [rng.uniform_int(N)]
    rng.uniform_int(N)
    bc = [0.0] * N
    seen, rounds, attempts, stopped = set(), [], 0, False
    for ui in range(N):
        if all((bc[a] if a < N else 0.0) / 2.0 >= c * N for a in checks):
            stopped, attempts = True, ui
            break
        v = rng.uniform_int(N)
        while v < N and v in seen:
            v = rng.uniform_int(N)
        if v >= N:
            continue
        seen.add(v)
        rounds.append(v)
        if v != 8:
            bc[8] += (N - 2) * N
    pv, _, obc, rs = oracle.run(g, N, c=c, x=1, z=0)
    assert stopped and rs["stopped"] and rs["attempts"] == attempts
    assert list(pv) == rounds and obc[8] == bc[8]


# --------------------------------------------------------------------------
# wrapped high-degree-neighbourhood prefix
# --------------------------------------------------------------------------

HUB_DEG = [50000, 40000, 70000, 62000]  # deg^2: 2.5e9, 1.6e9, 4.9e9 (wraps), 3.844e9


def hub_graph():
    """Hubs 0..3 with private leaves (ids 4 ...), hub h owning HUB_DEG[h]."""
    edges, nxt = [], len(HUB_DEG)
    for h, d in enumerate(HUB_DEG):
        edges.extend((h, nxt + j) for j in range(d))
        nxt += d
    return Graph.from_edges(nxt, np.array(edges, dtype=np.uint64))


def hdn_expect(g):
    deg = [int(x) for x in g.degrees()]
    prefix, run = [], 0
    for d in deg:
        run = (run + d * d) & M32    # (LOCINT)pow(deg, 2), summed as uint32
        prefix.append(run)
    return deg, prefix


def test_hdn_prefix_wraps_and_is_not_monotone():
    g = hub_graph()
    deg, prefix = hdn_expect(g)
    assert deg[:4] == HUB_DEG
    assert prefix[:4] == [2500000000, 4100000000, (4100000000 + 4900000000) & M32, 4254065408]
    assert prefix[2] < prefix[0] < prefix[1] < prefix[3]
    assert prefix[-1] == 4254065408 + sum(HUB_DEG)


@pytest.mark.parametrize("r_case", ["zero", "first", "peak", "peak+1", "after_wrap", "leaf", "last", "none"])
def test_hdn_draw_over_wrapped_prefix(r_case):
    g = hub_graph()
    deg, prefix = hdn_expect(g)
    N = g.N
    S = prefix[-1]
    r = {"zero": 0, "first": 2500000000, "peak": 4100000000, "peak+1": 4100000001,
         "after_wrap": 4254065408, "leaf": 4254065408 + 1234, "last": S, "none": S + 1}[r_case]
    # linear first match over the wrapped prefix, N - 1 when none
    u = next((i for i in range(N) if prefix[i] >= r), N - 1)
    want_u = {"zero": 0, "first": 0, "peak": 1, "peak+1": 3, "after_wrap": 3,
              "leaf": 4 + 1233, "last": N - 1, "none": N - 1}[r_case]
    assert u == want_u
    oracle.lib().oracle_srand48(ctypes.c_long(977))
    rng = Lrand48(977)
    adj_u = g.row[g.col[u]:g.col[u + 1]]
    v = int(adj_u[0 if deg[u] == 1 else rng.uniform_int(deg[u] - 1)])
    sdn_v = sum(deg[w] for w in g.row[g.col[v]:g.col[v + 1]]) & M32
    p = ctypes.c_double()
    got = oracle.lib().oracle_hdn_sample(N, g.col, g.row, r, ctypes.byref(p))
    assert got == v
    assert p.value == sdn_v / S
    # hub 2 (wrapped deg^2) is never the first match although its weight is not 0
    assert u != 2


# --------------------------------------------------------------------------
# draw-only mode
# --------------------------------------------------------------------------

@pytest.mark.parametrize("z", [0, 1, 2])
def test_draws_equal_the_loop_without_stop(z):
    g = Graph.kronecker(10, 16, seed=(1, 2), heuristic=1)
    pv, pr, _, rs = oracle.run(g, 300, c=0, x=7, z=z)
    dv, dp, ds = oracle.draws(g, 300, x=7, z=z)
    assert list(dv) == list(pv) and list(dp) == list(pr)
    assert ds["skipped"] == rs["skipped"] and ds["attempts"] == rs["attempts"]


def test_draws_refuse_dynamic_distributions():
    g = Graph.kronecker(8, 8, seed=(1, 2))
    for z in (3, 4, 5, 6):
        with pytest.raises(ValueError):
            oracle.draws(g, 10, z=z)
