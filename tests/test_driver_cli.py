"""The command-line driver's host-side behaviour, on the CPU: option
validation (src/axbc2d.c:2333-2459), the -d dump of the generated graph
(dump_rmat, 550-576, called at 2556-2559), the -a degree statistics
(analyze_deg, 1061-1115) and the -f edge-list reader (read_graph, 356-445).
Every run either fails validation or ends at -G (write the CSC as an edge
list and exit) before any device call."""
import glob
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "axbc_amd", "bin", "axbc2d")
if os.environ.get("AXBC_SAN_DIR") and os.path.exists(os.path.join(os.environ["AXBC_SAN_DIR"], "axbc2d")):
    DRIVER = os.path.join(os.environ["AXBC_SAN_DIR"], "axbc2d")  # the sanitizer build
GOLDEN = os.path.join(ROOT, "tests", "golden", "generator.json")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")


def drive(args, cwd):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=120, cwd=cwd)


# ---- validation --------------------------------------------------------------
@pytest.mark.parametrize("args,msg", [
    (["-p", "2x1", "-S", "8"], "R and C must be 1"),
    (["-p", "1x2", "-S", "8"], "R and C must be 1"),
    (["-p", "x", "-S", "8"], "Invalid proc mesh field"),
    (["-p", "1x1", "-S", "8", "-f", "g.txt"], "Unexpected option -f!"),
    (["-p", "1x1", "-f", "g.txt", "-n", "8", "-S", "8"], "Unexpected option -S!"),
    (["-p", "1x1", "-n", "8", "-E", "8"], "Unexpected option -E!"),
    # -f alone leaves N at 0: the check vertex count (default 1) fails first,
    # as in the reference's order of checks
    (["-p", "1x1", "-f", "g.txt"], "Number of vertices required is lager than N"),
    (["-p", "1x1", "-f", "g.txt", "-x", "0"], "Graph file (-f) and number of vertices (-n) must be specified"),
    (["-p", "1x1", "-n", "8"], "Graph file (-f) and number of vertices (-n) must be specified"),
    (["-p", "1x1", "-S", "8", "-z", "7"], "Select a strategy between 0 and 6"),
    (["-p", "1x1", "-S", "8", "-z", "-1"], "Select a strategy between 0 and 6"),
    (["-p", "1x1", "-S", "4", "-x", "17"], "Number of vertices required is lager than N"),
    (["-p", "1x1", "-S", "8", "-H", "2"], "H2 and H3 are not implemented"),
    (["-p", "1x1", "-S", "8", "-H", "one"], "Invalid Heuristic Option (-H): one"),
    (["-p", "1x1", "-S", "8", "-R", "4"], "Invalid random option (-R)"),
    (["-p", "1x1", "-S", "8", "-r", "256"], "Invalid start vertex"),
    (["-p", "1x1", "-S", "eight"], "Invalid scale (-S): eight"),
    (["-p", "1x1", "-S", "8", "-E", "x"], "Invalid edge factor"),
    (["-p", "1x1", "-S", "8", "-N", "many"], "Invalid number of bfs (-N): many"),
    (["-p", "1x1", "-f", "g.txt", "-n", "ten"], "Invalid number of vertices (-n): ten"),
    (["-p", "1x1", "-S", "8", "-r", "a"], "Invalid root vertex (-r): a"),
    (["-p", "1x1", "-S", "8", "-m"], "-m selects the reference's 2D multi-GPU path"),
    (["-p", "1x1", "-S", "8", "-B", "32"], "-B must be 0, 64, 128 or 256"),
    (["-p", "1x1", "-S", "8", "-B", "64", "-z", "4"], "-B batches need a static distribution"),
    (["-p", "1x1", "-S", "8", "-h"], "Usage:"),
    (["-p", "1x1", "-S", "8", "-Y"], "Usage:"),
])
def test_invalid_command_lines_fail(tmp_path, args, msg):
    r = drive(args, tmp_path)
    assert r.returncode != 0
    assert msg in r.stderr, r.stderr


def test_no_arguments_prints_usage(tmp_path):
    r = drive([], tmp_path)
    assert r.returncode != 0 and "Usage:" in r.stderr


def test_missing_graph_file_fails(tmp_path):
    r = drive(["-p", "1x1", "-f", str(tmp_path / "none.txt"), "-n", "8", "-G", "/dev/null"], tmp_path)
    assert r.returncode != 0 and "Cannot open file" in r.stderr


# ---- -d ------------------------------------------------------------------------
def read_pairs(path):
    a = np.loadtxt(path, dtype=np.uint64, ndmin=2)
    return a.reshape(-1, 2)


def golden_case(scale, nedges, s1, s2):
    with open(GOLDEN) as f:
        cases = json.load(f)["cases"]
    (c,) = [c for c in cases if (c["scale"], c["nedges"], c["seed1"], c["seed2"]) == (scale, nedges, s1, s2)]
    return c


def sha_of(pairs):
    k = np.sort((pairs[:, 0] << np.uint64(32)) | pairs[:, 1])
    return hashlib.sha256(k.astype("<u8").tobytes()).hexdigest()


@pytest.mark.parametrize("scale,ef,seed", [(12, 16, None), (8, 8, None), (10, 16, 1)])
def test_dump_is_the_mirrored_generator_output(tmp_path, scale, ef, seed):
    """-d at -H 0: the reference generator's edges (pinned by the golden
    digests of its own make_graph) and their reverses, self-loops once,
    duplicates kept, sorted by (u, v)."""
    args = ["-p", "1x1", "-S", str(scale), "-E", str(ef), "-d", "-G", "/dev/null"]
    if seed is not None:
        args += ["-s", str(seed)]
    r = drive(args, tmp_path)
    assert r.returncode == 0, r.stderr
    fname = f"rmat_S{scale}_EF{ef}.txt"
    assert f"DUMP THE GENERATED RMAT FILE IN {fname}" in r.stdout
    text = (tmp_path / fname).read_text()
    assert text.splitlines()[0].count("\t") == 1
    e = read_pairs(tmp_path / fname)
    s1, s2 = (23, 24) if seed is None else (seed, seed + 1)
    g = golden_case(scale, (1 << scale) * ef, s1, s2)
    assert e.shape[0] == 2 * g["count"] - g["self_loops"]
    key = (e[:, 0] << np.uint64(32)) | e[:, 1]
    assert np.all(np.diff(key.astype(np.uint64)) >= 0), "not sorted by (u, v)"
    fwd = e[e[:, 0] <= e[:, 1]]  # the generator's edges are (min, max)-oriented
    assert fwd.shape[0] == g["count"]
    assert sha_of(fwd) == g["sha256"]
    rev = e[e[:, 0] > e[:, 1]][:, ::-1]
    nl = fwd[fwd[:, 0] != fwd[:, 1]]
    assert sha_of(rev) == sha_of(nl)


def test_dump_normalised_with_heuristic(tmp_path):
    """-d at -H 1: normalised before the 1-degree reduction: the distinct
    non-loop edges of the -H 0 dump."""
    r0 = drive(["-p", "1x1", "-S", "10", "-E", "16", "-s", "1", "-d", "-G", "/dev/null"], tmp_path)
    assert r0.returncode == 0, r0.stderr
    e0 = read_pairs(tmp_path / "rmat_S10_EF16.txt")
    r1 = drive(["-p", "1x1", "-S", "10", "-E", "16", "-s", "1", "-H", "1", "-d", "-G", "/dev/null"], tmp_path)
    assert r1.returncode == 0, r1.stderr
    e1 = read_pairs(tmp_path / "rmat_S10_EF16.txt")
    k0 = (e0[:, 0] << np.uint64(32)) | e0[:, 1]
    k1 = (e1[:, 0] << np.uint64(32)) | e1[:, 1]
    assert np.all(np.diff(k1) > 0)
    expect = np.unique(k0[e0[:, 0] != e0[:, 1]])
    # the first sorted edge is kept even when it is a loop; here it is not one
    assert e0[0, 0] != e0[0, 1]
    assert np.array_equal(k1, expect)


def test_dump_round_trips_through_the_reader(tmp_path):
    """The graph read back from its dump (-f, -H 1) has the CSC of the
    generated graph with -H 1."""
    a = tmp_path / "a.txt"
    b = tmp_path / "b.txt"
    r = drive(["-p", "1x1", "-S", "10", "-E", "16", "-s", "2", "-H", "1", "-d", "-G", str(a)], tmp_path)
    assert r.returncode == 0, r.stderr
    r = drive(["-p", "1x1", "-f", str(tmp_path / "rmat_S10_EF16.txt"), "-n", "1024", "-H", "1", "-G", str(b)],
              tmp_path)
    assert r.returncode == 0, r.stderr
    assert a.read_text() == b.read_text()


def test_dump_only_for_generated_graphs(tmp_path):
    src = tmp_path / "g.txt"
    src.write_text("0 1\n1 2\n")
    r = drive(["-p", "1x1", "-f", str(src), "-n", "4", "-d", "-G", "/dev/null"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert not glob.glob(str(tmp_path / "rmat_*"))


# ---- -a ------------------------------------------------------------------------
def test_degree_stats_by_hand(tmp_path):
    """Degrees 150 (a hub), 12 (a second hub), 2, 1 and 0: the 16 bins with
    lower limits 0..10, 100, 1000, 1e4, 1e5, 1e7 count one vertex of degree
    150 in bin 11, one of degree 12 in bin 10, and so on."""
    lines = [f"0 {v}" for v in range(1, 151)]          # hub 0: degree 150
    lines += [f"151 {v}" for v in range(152, 164)]     # hub 151: degree 12
    lines += ["164 165", "165 166"]                   # path: degrees 1, 2, 1
    n = 170                                           # 167..169 isolated
    src = tmp_path / "g.txt"
    src.write_text("\n".join(lines) + "\n")
    r = drive(["-p", "1x1", "-f", str(src), "-n", str(n), "-a", "-G", "/dev/null"], tmp_path)
    assert r.returncode == 0, r.stderr
    (f,) = glob.glob(str(tmp_path / "degree_stats_*_0"))
    bins = [0] * 16
    bins[0] = 3                # 167, 168, 169
    bins[1] = 150 + 12 + 2     # leaves of both hubs, path ends
    bins[2] = 1                # path middle
    bins[10] = 1               # degree 12 in [10, 100)
    bins[11] = 1               # degree 150 in [100, 1000)
    assert open(f).read() == "".join(f" {b}," for b in bins) + "\n"


def test_degree_stats_append_and_count_every_vertex(tmp_path):
    r = drive(["-p", "1x1", "-S", "10", "-E", "16", "-s", "1", "-H", "1", "-a", "-G", str(tmp_path / "g.txt")],
              tmp_path)
    assert r.returncode == 0, r.stderr
    (f,) = glob.glob(str(tmp_path / "degree_stats_*_0"))
    row = [int(x) for x in open(f).read().strip().rstrip(",").split(",")]
    assert len(row) == 16 and sum(row) == 1024
    col = read_pairs(tmp_path / "g.txt")[:, 1].astype(np.int64)
    deg = np.bincount(col, minlength=1024)
    assert row[0] == int((deg == 0).sum()) and row[1] == int((deg == 1).sum())
    assert sum(row[10:]) == int((deg >= 10).sum())


# ---- -f ------------------------------------------------------------------------
def test_reader_comments_and_header(tmp_path):
    """Lines containing '#' anywhere are comments; the "# Nodes: n Edges: m"
    header is echoed as "N=n E=m"; pairs may be tab- or space-separated;
    the list is mirrored unless -U."""
    src = tmp_path / "g.txt"
    src.write_text("# Directed graph (each unordered pair of nodes is saved once)\n"
                   "# Nodes: 6 Edges: 5\n"
                   "# FromNodeId\tToNodeId\n"
                   "0\t1\n"
                   "1 2\n"
                   "2\t3 # a comment after a pair makes the line a comment\n"
                   "3 4\n"
                   "\n"
                   "4\t5\n"
                   "5 5\n")
    out = tmp_path / "csc.txt"
    r = drive(["-p", "1x1", "-f", str(src), "-n", "6", "-G", str(out)], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "N=6 E=5" in r.stdout
    assert "Edges read from file (after mirror): 9" in r.stdout   # 5 pairs + 4 reverses (no loop reverse)
    csc = [tuple(map(int, l.split("\t"))) for l in out.read_text().splitlines()]
    # CSC lines are "row<TAB>column" by column; the self-loop and the
    # commented pair are gone
    edges = {(0, 1), (1, 2), (3, 4), (4, 5)}
    expect = sorted({(u, v) for a, b in edges for u, v in ((a, b), (b, a))}, key=lambda t: (t[1], t[0]))
    assert csc == expect
    r = drive(["-p", "1x1", "-f", str(src), "-n", "6", "-U", "-G", str(out)], tmp_path)
    assert r.returncode == 0, r.stderr
    csc = [tuple(map(int, l.split("\t"))) for l in out.read_text().splitlines()]
    assert csc == sorted(edges, key=lambda t: (t[1], t[0]))


def test_reader_rejects_out_of_range_ids(tmp_path):
    src = tmp_path / "g.txt"
    src.write_text("# Nodes: 4 Edges: 2\n0 1\n1 4\n")
    r = drive(["-p", "1x1", "-f", str(src), "-n", "4", "-G", "/dev/null"], tmp_path)
    assert r.returncode != 0
    assert "line 2 found invalid edge for N=4: (1, 4)" in r.stderr
