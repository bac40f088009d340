"""Whole pivot loop (sampler + stop test + rounds) on the device against
the oracle's loop: same pivot sequence, same round/skip counts, same stop
decision, bc within 1e-9 relative -- one root per round and batched."""
import numpy as np
import pytest

from axbc_amd import Engine, Graph
from axbc_amd.pivots import PivotLoop
from oracle import oracle
from tests.test_gpu_plain import close

pytestmark = pytest.mark.gpu


def test_require(require_gpu):
    pass


@pytest.mark.parametrize("z,c,x,batch,heur", [
    (0, 0, 1, 0, 0), (0, 1, 1, 0, 1), (1, 0, 4, 0, 1), (3, 0, 1, 0, 0), (4, 0, 1, 0, 1), (5, 0, 1, 0, 0),
    (0, 0, 1, 64, 0), (1, 1, 8, 64, 1), (1, 0, 1, 128, 1), (0, 2, 3, 64, 0),
])
def test_loop_matches_oracle(require_gpu, z, c, x, batch, heur):
    g = Graph.kronecker(11, 16, seed=(1, 2), heuristic=heur)
    nbfs = 300
    print(f"[loop z{z} c{c} x{x} B{batch} H{heur}]", flush=True)
    pv, pr, bc_ref, rs = oracle.run(g, nbfs, c=c, x=x, z=z)
    with Engine(g) as e:
        res = PivotLoop(e, nbfs, c=c, x=x, z=z, batch=batch).run()
        assert res.pivots == [int(v) for v in pv]
        np.testing.assert_array_equal(np.array(res.probs), pr)
        assert res.nrounds == rs["nrounds"] and res.skipped == rs["skipped"]
        assert res.stopped == rs["stopped"]
        ok, err = close(e.bc(), bc_ref)
        assert ok, err
