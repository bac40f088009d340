"""Plain one-root-per-round HIP path against the CPU oracle.

Per round: lvl and sigma bit-exact (int32 / uint32 with wrap-around), delta
and the accumulated bc within 1e-9 relative (FP64 on both sides; the device
sums dependencies in a different order).
"""
import numpy as np
import pytest

from axbc_amd import Engine, Graph
from oracle import oracle
from tests.parity import close

pytestmark = pytest.mark.gpu
def check_rounds(g, roots, label, pull=None):
    print(f"[{label}] N={g.N} ned={g.ned} roots={len(roots)} pull={pull}", flush=True)
    bc_ref = np.zeros(g.N)
    with Engine(g) as e:
        if pull is not None:
            from axbc_amd import _ffi
            assert _ffi.hip().axbc_set_plain_pull(e.ctx, pull) == 0
        for k, v0 in enumerate(roots):
            p = 1.0 / g.N
            st = e.pivot(int(v0), p, stats=True)
            lvl, sig, dl, bc_ref, ost = oracle.pivot(g, int(v0), p, bc_ref)
            assert np.array_equal(e.lvl(), lvl), f"{label} round {k} root {v0}: lvl differs"
            assert np.array_equal(e.sigma(), sig), f"{label} round {k} root {v0}: sigma differs"
            ok, err = close(e.delta(), dl)
            assert ok, f"{label} round {k} root {v0}: delta rel err {err}"
            ok, err = close(e.bc(), bc_ref)
            assert ok, f"{label} round {k} root {v0}: bc rel err {err}"
            for key in ("nvisited", "nlevels", "e_r", "e_dag", "all"):
                assert st[key] == ost[key], f"{label} round {k}: {key} {st[key]} != {ost[key]}"


def test_require(require_gpu):
    pass


# pull ratio (x16) of the plain levels: 0 = every level top-down, 32 = the
# default switch, 65536 = bottom-up from level 2 on (symmetric graphs only;
# the directed graph keeps every level top-down whatever the ratio)
@pytest.mark.parametrize("pull", [0, 32, 65536])
@pytest.mark.parametrize("scale,ef,seed,undirected,heur", [
    (10, 16, (1, 2), True, 0),
    (12, 16, (23, 24), True, 1),
    (12, 8, (3, 4), False, 0),
])
def test_kronecker_rounds(require_gpu, scale, ef, seed, undirected, heur, pull):
    g = Graph.kronecker(scale, ef, seed=seed, undirected=undirected, heuristic=heur)
    deg = g.degrees()
    rng = np.random.default_rng(scale * 7 + heur)
    cand = np.flatnonzero(deg > 0)
    roots = list(rng.choice(cand, size=8, replace=False)) + [int(np.argmax(deg))]
    check_rounds(g, roots, f"kron s{scale} ef{ef} H{heur} U{int(not undirected)}", pull=pull)


@pytest.mark.parametrize("pull", [0, 65536])
def test_hub_paths(require_gpu, pull):
    # star + clique core: the centre exceeds the hub degree threshold, so
    # both forward and backward hub passes run (bottom-up: the centre's
    # list is dealt in hub chunks while it is not yet found)
    n = 6000
    edges = [(0, i) for i in range(1, n)] + [(i, i + 1) for i in range(1, n - 1)]
    g = Graph.from_edges(n, np.array(edges), undirected=True)
    check_rounds(g, [0, 1, 2999, n - 1], "star-path", pull=pull)


@pytest.mark.parametrize("pull", [0, 65536])
def test_hub_found_bottom_up(require_gpu, pull):
    """A hub of 5000 leaves three steps from the root: with bottom-up levels
    from level 2 on, the hub is found while its list is dealt in hub chunks
    (the chunk that sees its predecessor claims it, every chunk adds its
    path counts), and its leaves one level later."""
    n = 5010
    hub = 5
    edges = [(0, 1), (1, 2), (2, 3), (3, hub), (0, 4), (4, 3)] + [(hub, i) for i in range(10, n)]
    edges += [(i, i + 1) for i in range(10, n - 1, 2)]  # leaf pairs: two paths into each from the hub
    g = Graph.from_edges(n, np.array(edges), undirected=True)
    check_rounds(g, [0, 1, 10, n - 1], "path into a hub", pull=pull)


def test_path_depths(require_gpu):
    n = 3000
    g = Graph.from_edges(n, np.array([(i, i + 1) for i in range(n - 1)]), undirected=True)
    check_rounds(g, [0, n // 2, n - 1], "long path")


def layered(a: int, layers: int) -> Graph:
    """Root 0, then `layers` layers of a vertices, consecutive layers
    completely connected: sigma of layer j (from 0) is a^j mod 2^32."""
    e = [(0, 1 + x) for x in range(a)]
    for j in range(layers - 1):
        xs = np.arange(a)
        u = np.repeat(1 + j * a + xs, a)
        w = np.tile(1 + (j + 1) * a + xs, a)
        e.extend(zip(u.tolist(), w.tolist()))
    return Graph.from_edges(1 + layers * a, np.array(e))


def test_sigma_wraps_to_zero(require_gpu):
    # a = 256: sigma of layer 4 is 2^32 = 0 mod 2^32, so x = (delta+1)/0 is
    # infinite and the dependencies above it become inf/NaN, as in the
    # reference's uint32 arithmetic; both sides must agree position by position
    g = layered(256, 6)
    check_rounds(g, [0, 1 + 5 * 256 + 7, 1 + 2 * 256 + 3], "sigma wraps to 0")


@pytest.mark.parametrize("undirected", [True, False])
def test_plain_config_per_round(require_gpu, undirected):
    """The CPU-runnable plain configuration, round by round: Kronecker SCALE
    16, edge factor 16, seed pair (1, 2), -z 0 -N 256 (once symmetrised and
    once with -U), the oracle's pivot sequence; every round's lvl and sigma
    bit-exact, delta and bc within 1e-9."""
    g = Graph.kronecker(16, 16, seed=(1, 2), undirected=undirected, heuristic=0)
    pv, pr, bc, rs = oracle.run(g, 256, c=0, x=1, z=0)
    assert rs["nrounds"] == len(pv) > 150  # isolated vertices drawn are skipped
    check_rounds(g, pv, f"scale16-{'sym' if undirected else 'dir'}")
