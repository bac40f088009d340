"""Graph ingest (mirror, norm_graph, degree reduction, CSC) against a
straightforward Python restatement of src/axbc2d.c:333-351, 695-841,
866-907 on small random edge lists, and properties of the lattice."""
import numpy as np
import pytest

from axbc_amd import Graph


def py_mirror(e):
    return list(e) + [(v, u) for (u, v) in e if u != v]


def py_norm(e):
    e = sorted(e)
    if not e:
        return []
    out = [e[0]]
    for n in range(1, len(e)):
        if e[n] != e[n - 1] and e[n][0] != e[n][1]:
            out.append(e[n])
    return out


def py_degree_reduction(e, N):
    e = sorted(e)
    kept, gone = [], []
    prev_u = None
    for i in range(len(e) - 1):
        (u, v), (nu, nv) = e[i], e[i + 1]
        if u == v or (u == nu and v == nv):
            prev_u = u
            continue
        if u != nu and u != prev_u:
            gone.append((v, u))
        else:
            kept.append((u, v))
        prev_u = u
    u, v = e[-1]
    if u == prev_u:
        kept.append((u, v))
    elif u != v:
        gone.append((v, u))
    gone.sort()
    at = lambda k: gone[k] if k < len(gone) else (0, 0)  # noqa: E731
    pnod, out = 0, []
    if gone:
        for x in kept:
            while x[0] > at(pnod)[0] and pnod < len(gone):
                pnod += 1
            if x == at(pnod):
                pnod += 1
                continue
            out.append(x)
    else:
        out = kept
    reach = [0] * N
    for i in range(min(pnod, len(gone))):
        reach[gone[i][0]] += 1
    return out, reach, pnod


def py_csc(e, N):
    col = [0] * (N + 1)
    for (_, v) in e:
        col[v + 1] += 1
    for j in range(N):
        col[j + 1] += col[j]
    fill = col[:N]
    row = [0] * len(e)
    fill = list(fill)
    for (u, v) in e:
        row[fill[v]] = u
        fill[v] += 1
    return col, row


@pytest.mark.parametrize("seed", range(20))
@pytest.mark.parametrize("heur", [0, 1])
def test_pipeline_matches_restatement(seed, heur):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(4, 40))
    m = int(rng.integers(1, 3 * N))
    e = [tuple(int(x) for x in r) for r in rng.integers(0, N, size=(m, 2))]
    if seed % 3 == 0:  # pendant-rich trees
        e = [(int(rng.integers(0, i)), i) for i in range(1, N)]
    g = Graph.from_edges(N, np.array(e), undirected=True, heuristic=heur)
    x = py_mirror(e)
    reach = [0] * N
    rem = 0
    if heur:
        x = py_norm(x)
        x, reach, rem = py_degree_reduction(x, N)
    x = py_norm(x)
    col, row = py_csc(x, N)
    assert list(g.col) == col
    assert list(g.row) == row
    assert list(g.reach) == reach
    assert g.rem_ed == rem


def test_first_sorted_self_loop_survives_norm():
    # norm_graph keeps ed[0] unconditionally: a loop on the smallest vertex
    g = Graph.from_edges(4, np.array([(0, 0), (1, 2), (2, 3)]), undirected=True)
    assert list(g.row[g.col[0]:g.col[1]]) == [0]


def test_undirected_flag_off_keeps_direction():
    g = Graph.from_edges(3, np.array([(0, 1), (1, 2)]), undirected=False)
    # column = second endpoint: adjacency of 1 is {0}, of 2 is {1}
    assert list(g.col) == [0, 0, 1, 2] and list(g.row) == [0, 1]


def test_lattice_properties():
    g = Graph.lattice(64, 0.03, 3)
    assert g.N == 64 * 64
    deg = g.degrees()
    assert deg.max() <= 4
    # undirected: every edge appears in both adjacencies
    e = set()
    for u in range(g.N):
        for w in g.row[g.col[u]:g.col[u + 1]]:
            e.add((u, int(w)))
            assert abs(int(w) - u) in (1, 64)
    assert all((w, u) in e for (u, w) in e)
    # one component: BFS from any non-isolated vertex reaches all of them
    start = int(np.flatnonzero(deg)[0])
    seen = {start}
    stack = [start]
    while stack:
        u = stack.pop()
        for w in g.row[g.col[u]:g.col[u + 1]]:
            if int(w) not in seen:
                seen.add(int(w))
                stack.append(int(w))
    assert len(seen) == int((deg > 0).sum())
    full = 2 * 64 * 63
    assert 0.9 * full < len(e) / 2 <= full


def test_serving_footprint_and_workload():
    from axbc_amd.serve import footprint_bytes, poisson_jobs
    assert footprint_bytes(1 << 20, 30_000_000) > 4 * 30_000_000
    assert footprint_bytes(1 << 20, 1000, batch=64) > footprint_bytes(1 << 20, 1000)
    js = poisson_jobs(200, 20.0, seed=1)
    t = np.array([a for a, _ in js])
    assert (np.diff(t) > 0).all() and 5 < t[-1] < 20  # ~10 s of arrivals at 20/s
    s = np.array([j.scale for _, j in js])
    b = np.array([j.budget for _, j in js])
    assert s.min() >= 12 and s.max() <= 20 and b.min() >= 64 and b.max() <= 2048
    assert {j.edgefactor for _, j in js} <= {8, 16, 32}
