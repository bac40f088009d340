"""The C driver (axbc_amd/bin/axbc2d, the reference's command line) against
the oracle's pivot loop: same pivots in the per-pivot log lines, same
round/skip counts, -o result file within 1e-9 relative."""
import os
import re
import subprocess

import numpy as np
import pytest

from axbc_amd import Graph
from oracle import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "axbc_amd", "bin", "axbc2d")


def test_require(require_gpu):
    pass


@pytest.mark.parametrize("flags", [
    ["-z", "0", "-c", "1", "-x", "1"],
    ["-z", "1", "-H", "1", "-c", "1", "-x", "16"],
    ["-z", "1", "-H", "1", "-c", "0", "-B", "64"],
    ["-z", "4", "-c", "0"],
    ["-z", "6", "-H", "1", "-c", "0"],
    ["-z", "1", "-H", "1", "-c", "0", "-B", "64", "-g"],  # preprocessing on the device
])
def test_driver_matches_oracle(require_gpu, tmp_path, flags):
    scale, nbfs = 10, 200
    out = tmp_path / "res"
    cmd = [DRIVER, "-p", "1x1", "-S", str(scale), "-E", "16", "-s", "1", "-N", str(nbfs), "-o", str(out)] + flags
    print("[driver]", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    piv = [int(m.group(1)) for m in re.finditer(r"^Perform \d+/\d+ BC from (\d+) ", r.stdout, re.M)]
    rounds = int(re.search(r"task 0 BC rounds: (\d+)", r.stdout).group(1))
    skipped = int(re.search(r"task 0 BC skipped: (\d+)", r.stdout).group(1))
    get = lambda f: flags[flags.index(f) + 1] if f in flags else None  # noqa: E731
    heur = int(get("-H") or 0)
    g = Graph.kronecker(scale, 16, seed=(1, 2), heuristic=heur)
    pv, pr, bc, rs = oracle.run(g, nbfs, c=int(get("-c") or 1), x=int(get("-x") or 1), z=int(get("-z")))
    assert rounds == rs["nrounds"] and skipped == rs["skipped"]
    assert piv[:rounds] == [int(v) for v in pv]
    res = np.loadtxt(str(out) + "_1X1_0.log")
    assert res.shape == (g.N, 2) and (res[:, 0] == np.arange(g.N)).all()
    exp = bc / 2.0 / max(rounds, 1)
    assert np.allclose(res[:, 1], exp, rtol=1e-9, atol=0.0051)


@pytest.mark.parametrize("directed", [False, True])
def test_plain_config_scale16(require_gpu, tmp_path, directed):
    """The CPU-runnable plain configuration: SCALE 16, edge factor 16, seed
    pair (1, 2), -z 0 -N 256 with the default adaptive stop (c=1, x=1); once
    symmetrised and once with -U (which turns mirroring off)."""
    out = tmp_path / "res"
    cmd = [DRIVER, "-p", "1x1", "-S", "16", "-E", "16", "-s", "1", "-z", "0", "-N", "256", "-q", "-o", str(out)]
    if directed:
        cmd.append("-U")
    print("[driver]", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    rounds = int(re.search(r"task 0 BC rounds: (\d+)", r.stdout).group(1))
    skipped = int(re.search(r"task 0 BC skipped: (\d+)", r.stdout).group(1))
    g = Graph.kronecker(16, 16, seed=(1, 2), undirected=not directed, heuristic=0)
    pv, pr, bc, rs = oracle.run(g, 256, c=1, x=1, z=0)
    assert rounds == rs["nrounds"] and skipped == rs["skipped"]
    res = np.loadtxt(str(out) + "_1X1_0.log")
    exp = bc / 2.0 / max(rounds, 1)
    assert np.allclose(res[:, 1], exp, rtol=1e-9, atol=0.0051)
