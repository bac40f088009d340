"""The reference's own device call sequence (bc_func_mono + main) through
libaxbc_legacy.so, as AxBC's host code would link it, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from axbc_amd import Graph
from oracle import oracle
from tests.parity import close

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "legacy", "bin", "legacy_host")


def test_require(require_gpu):
    pass


@pytest.mark.parametrize("heur", [0, 1])
def test_legacy_sequence(require_gpu, tmp_path, heur):
    assert os.path.exists(BIN), "legacy test host not built (python -m axbc_amd.build)"
    g = Graph.kronecker(11, 16, seed=(5, 6), heuristic=heur)
    gp = tmp_path / "g.bin"
    with open(gp, "wb") as f:
        np.array([g.N, g.ned], np.uint64).tofile(f)
        np.array([heur], np.int32).tofile(f)
        g.col.astype(np.uint32).tofile(f)
        g.row.astype(np.uint32).tofile(f)
        g.reach.astype(np.uint32).tofile(f)
    deg = g.degrees()
    roots = np.flatnonzero(deg > 0)[::37][:24].astype(np.uint32)
    probs = np.linspace(1e-4, 1e-3, len(roots))
    pp = tmp_path / "p.bin"
    with open(pp, "wb") as f:
        np.array([len(roots)], np.uint64).tofile(f)
        for v, p in zip(roots, probs):
            np.array([v], np.uint32).tofile(f)
            np.array([p], np.float64).tofile(f)
    out = tmp_path / "bc.bin"
    print(f"[legacy H{heur}] N={g.N} roots={len(roots)}", flush=True)
    r = subprocess.run([BIN, str(gp), str(pp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    bc = np.fromfile(out, np.float64)
    bc_ref = np.zeros(g.N)
    for v, p in zip(roots, probs):
        _, _, _, bc_ref, _ = oracle.pivot(g, int(v), float(p), bc_ref)
    ok, err = close(bc, bc_ref)
    assert ok, err
