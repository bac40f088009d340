"""Kronecker generation on the device (axbc_kronecker_generate) against the
reference Graph500 generator's fixtures (tests/golden/generator.json) and
against the host generator element for element; the whole graph built on
the device (axbc_kronecker_ingest) against the host pipeline bit for bit."""
import json
import os

import numpy as np
import pytest

from axbc_amd import Graph, kronecker_edge_list
from axbc_amd.graph import INITIATOR, kronecker_keys_device
from tests.golden.make_generator_fixtures import digest
from tests.test_gpu_ingest import same

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "generator.json")))


def test_require(require_gpu):
    pass


def pairs(keys: np.ndarray) -> np.ndarray:
    return np.stack([keys >> np.uint64(32), keys & np.uint64(0xFFFFFFFF)], axis=1).astype(np.int64)


@pytest.mark.parametrize("case", FIX["cases"], ids=lambda c: f"s{c['scale']}-m{c['nedges']}-{c['seed1']}")
def test_matches_reference_generator(require_gpu, case):
    assert tuple(FIX["initiator"]) == INITIATOR
    print(f"[gen] fixture s{case['scale']} m{case['nedges']} seed {case['seed1']}", flush=True)
    d = digest(pairs(kronecker_keys_device(case["scale"], case["nedges"], case["seed1"], case["seed2"])))
    for k in ("count", "sha256", "head", "tail", "self_loops"):
        assert d[k] == case[k], k


@pytest.mark.parametrize("scale,nedges,seed", [
    (1, 5, (1, 2)), (3, 1, (7, 8)), (4, 2, (3, 4)), (6, 21, (9, 10)), (8, 1000, (1, 2)),
    (10, 3, (11, 12)), (12, 65536, (23, 24)), (16, 1 << 20, (2, 3)), (18, 1 << 22, (5, 6)),
    (20, 1 << 24, (2, 3)),
])
def test_equals_host_generator(require_gpu, scale, nedges, seed):
    print(f"[gen] host vs device s{scale} m{nedges} seed {seed}", flush=True)
    host = kronecker_edge_list(scale, nedges, seed[0], seed[1]).astype(np.uint64)
    hk = (host[:, 0] << np.uint64(32)) | host[:, 1]
    dk = kronecker_keys_device(scale, nedges, seed[0], seed[1])
    assert dk.shape == hk.shape
    bad = np.flatnonzero(dk != hk)
    assert bad.size == 0, f"{bad.size} edges differ, first at {bad[:4]}"


def test_no_edges(require_gpu):
    assert kronecker_keys_device(5, 0, 1, 2).size == 0


@pytest.mark.parametrize("undirected", [True, False])
@pytest.mark.parametrize("heuristic", [0, 1])
def test_graph_on_device_equals_host(require_gpu, undirected, heuristic):
    print(f"[gen] graph s14 U={int(not undirected)} H={heuristic}", flush=True)
    host = Graph.kronecker(14, 16, seed=(23, 24), undirected=undirected, heuristic=heuristic)
    dev = Graph.kronecker_device(14, 16, seed=(23, 24), undirected=undirected, heuristic=heuristic)
    same(dev, host, "kron s14")
    assert dev.generated == host.generated


def test_graph_on_device_scale22(require_gpu):
    print("[gen] graph s22 H=1", flush=True)
    host = Graph.kronecker(22, 16, seed=(2, 3), heuristic=1)
    dev = Graph.kronecker_device(22, 16, seed=(2, 3), heuristic=1)
    same(dev, host, "kron s22")
