"""Graph construction on the device (axbc_ingest) against the host pipeline
(graph.c, the restatement of the reference's mirror / norm_graph / 1-degree
reduction / build_csc): col, row, reach, the removed-edge count and the
edge count bit for bit."""
import numpy as np
import pytest

from axbc_amd import Graph
from axbc_amd.graph import kronecker_edge_list

pytestmark = pytest.mark.gpu


def test_require(require_gpu):
    pass


def same(a: Graph, b: Graph, label: str):
    assert a.ned == b.ned, f"{label}: ned {a.ned} != {b.ned}"
    np.testing.assert_array_equal(a.col, b.col, err_msg=label)
    np.testing.assert_array_equal(a.row, b.row, err_msg=label)
    np.testing.assert_array_equal(a.reach, b.reach, err_msg=label)
    assert a.rem_ed == b.rem_ed, f"{label}: rem_ed {a.rem_ed} != {b.rem_ed}"


def check(N, pairs, undirected, heuristic, label):
    pairs = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    print(f"[ingest] {label} N={N} m={len(pairs)} U={int(not undirected)} H={heuristic}", flush=True)
    host = Graph.from_edges(N, pairs, undirected=undirected, heuristic=heuristic)
    keys = (pairs[:, 0] << np.uint64(32)) | pairs[:, 1]
    dev = Graph.from_keys_device(N, keys, undirected=undirected, heuristic=heuristic)
    same(dev, host, label)


@pytest.mark.parametrize("undirected", [True, False])
@pytest.mark.parametrize("heuristic", [0, 1])
@pytest.mark.parametrize("scale,seed", [(8, (1, 2)), (12, (23, 24)), (16, (2, 3))])
def test_kronecker(require_gpu, scale, seed, undirected, heuristic):
    e = kronecker_edge_list(scale, (1 << scale) * 16, seed[0], seed[1])
    check(1 << scale, e, undirected, heuristic, f"kron s{scale}")


@pytest.mark.parametrize("undirected", [True, False])
@pytest.mark.parametrize("heuristic", [0, 1])
def test_random_with_pendants_loops_duplicates(require_gpu, undirected, heuristic):
    rng = np.random.default_rng(11)
    for trial in range(6):
        N = int(rng.integers(2, 200))
        m = int(rng.integers(1, 600))
        # a sparse core, pendant edges, loops and repeated edges
        e = rng.integers(0, N, size=(m, 2))
        e = np.vstack([e, e[: m // 5], np.stack([e[:5, 0], e[:5, 0]], 1)])
        rng.shuffle(e)
        check(N, e, undirected, heuristic, f"random {trial}")


@pytest.mark.parametrize("heuristic", [0, 1])
def test_edge_cases(require_gpu, heuristic):
    check(4, np.zeros((0, 2)), True, heuristic, "empty")
    check(4, [(1, 2)], True, heuristic, "one edge")
    check(4, [(1, 2)], False, heuristic, "one directed edge")
    check(4, [(0, 0)], True, heuristic, "one loop")
    check(5, [(0, 0), (0, 1), (1, 2), (2, 3), (3, 4)], True, heuristic, "path with a loop at 0")
    check(6, [(0, 1), (2, 3), (4, 5)], True, heuristic, "isolated edges")
    check(6, [(0, 1), (0, 2), (0, 3), (3, 4), (4, 5)], False, heuristic, "directed star and tail")


def test_lattice_and_engine(require_gpu):
    from axbc_amd import Engine
    from oracle import oracle
    from tests.parity import close
    rng = np.random.default_rng(3)
    side = 64
    e = [(r * side + c, r * side + c + 1) for r in range(side) for c in range(side - 1)]
    e += [(r * side + c, (r + 1) * side + c) for r in range(side - 1) for c in range(side)]
    e = np.array(e)[rng.random(len(e)) > 0.03]
    check(side * side, e, True, 1, "lattice 64")
    # the device-built graph drives the engine like the host-built one
    keys = (e[:, 0].astype(np.uint64) << np.uint64(32)) | e[:, 1].astype(np.uint64)
    g = Graph.from_keys_device(side * side, keys, undirected=True, heuristic=1)
    bc_ref = np.zeros(g.N)
    v0 = int(np.flatnonzero(g.degrees() > 0)[0])
    _, _, _, bc_ref, _ = oracle.pivot(g, v0, 1.0 / g.N, bc_ref)
    with Engine(g) as eng:
        eng.pivot(v0, 1.0 / g.N)
        ok, err = close(eng.bc(), bc_ref)
    assert ok, err
