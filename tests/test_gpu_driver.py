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


@pytest.mark.parametrize("flags", [
    ["-z", "1", "-H", "1", "-c", "1", "-x", "8"],
    ["-z", "6", "-H", "0", "-c", "0"],
    ["-z", "0", "-H", "1", "-c", "0", "-B", "64", "-g"],
    ["-z", "0", "-H", "1", "-c", "0", "-B", "256"],  # two bit-packed sweeps side by side, applied in two halves
    ["-z", "1", "-H", "1", "-c", "1", "-x", "16", "-B", "256"],
])
def test_driver_on_a_file_matches_oracle(require_gpu, tmp_path, flags):
    """-f: a text edge list with '#' comment lines and a "# Nodes:" header,
    pairs tab-separated, read, mirrored and preprocessed (on the host, or
    on the device with -g); pivots, round and skip counts and the result
    file match the oracle's loop on the same edges."""
    rng = np.random.default_rng(11)
    n = 3000
    # a sparse random graph with hubs, a few isolated vertices and pendants
    m = 12000
    u = rng.integers(0, n - 50, m)
    v = (u + rng.zipf(1.6, m)) % (n - 50)
    hubs = rng.integers(0, 20, 400)
    leaves = rng.integers(0, n - 50, 400)
    e = np.concatenate([np.stack([u, v], 1), np.stack([hubs, leaves], 1),
                        np.array([[n - 40 + k, n - 20 + k] for k in range(10)])])
    src = tmp_path / "graph.txt"
    with open(src, "w") as f:
        f.write("# Undirected test graph\n# Nodes: %d Edges: %d\n# FromNodeId\tToNodeId\n" % (n, len(e)))
        for a, b in e:
            f.write(f"{a}\t{b}\n")
    out = tmp_path / "res"
    nbfs = 200
    cmd = [DRIVER, "-p", "1x1", "-f", str(src), "-n", str(n), "-N", str(nbfs), "-o", str(out)] + flags
    print("[driver -f]", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert f"N={n} E={len(e)}" in r.stdout
    piv = [int(x.group(1)) for x in re.finditer(r"^Perform \d+/\d+ BC from (\d+) ", r.stdout, re.M)]
    rounds = int(re.search(r"task 0 BC rounds: (\d+)", r.stdout).group(1))
    skipped = int(re.search(r"task 0 BC skipped: (\d+)", r.stdout).group(1))
    get = lambda fl: flags[flags.index(fl) + 1] if fl in flags else None  # noqa: E731
    g = Graph.from_edges(n, e, undirected=True, heuristic=int(get("-H")))
    pv, pr, bc, rs = oracle.run(g, nbfs, c=int(get("-c")), x=int(get("-x") or 1), z=int(get("-z")))
    assert rounds == rs["nrounds"] and skipped == rs["skipped"]
    assert piv[:rounds] == [int(x) for x in pv]
    res = np.loadtxt(str(out) + "_1X1_0.log")
    assert res.shape == (n, 2)
    exp = bc / 2.0 / max(rounds, 1)
    assert np.allclose(res[:, 1], exp, rtol=1e-9, atol=0.0051)
