"""The benchmark workload at its full size (Kronecker SCALE 22, edgefactor
16, seed pair (2, 3), -H 1, high-degree-neighbourhood roots of
srand48(4343543), batches of 128 at chunk 512), the 4096^2 lattice and
Kronecker SCALE 24: the root sequence against the oracle's draws, sampled
roots' levels and path counts bit-exact and delta within 1e-9 relative
against the oracle, and each batch's bc against the same roots run one per
round on the device (bc and delta are FP64)."""
import numpy as np
import pytest

from axbc_amd import Engine, Graph, _ffi
from axbc_amd.pivots import PivotLoop
from oracle import oracle
from tests.test_gpu_batch import delta_of
from tests.parity import close

pytestmark = pytest.mark.gpu


def test_require(require_gpu):
    pass


def test_scale22_batch_against_oracle_and_plain(require_gpu):
    """The benchmark loop itself (-z 1, c = 1, x = 1024, batches of 128 at
    chunk 512) for its first 1024 rounds: the root sequence, probabilities
    and skip count equal the oracle's draws (the stop test never holds here:
    check vertices removed by the 1-degree reduction keep bc = 0).  Then the
    last batch's 128 roots again with counters: three roots' levels, path
    counts (bit-exact), delta (1e-9) and counters against the oracle, and
    the batch's bc against the same roots run one per round."""
    print("[fullsize] building SCALE 22 on the device", flush=True)
    g = Graph.kronecker_device(22, 16, seed=(2, 3), heuristic=1)
    L = _ffi.hip()
    with Engine(g) as e:
        assert L.axbc_set_batch_width(e.ctx, 128) == 0
        assert L.axbc_set_batch_chunk(e.ctx, 512) == 0
        loop = PivotLoop(e, 4096, c=1, x=1024, z=1, batch=128)
        while loop.res.nrounds < 1024:
            assert loop.step() == 128
        res = loop.res
        assert not res.stopped
        print(f"[fullsize] 1024 rounds in {res.attempts} attempts; oracle draws", flush=True)
        dv, dp, ds = oracle.draws(g, res.attempts, x=1024, z=1)
        assert list(dv) == list(res.pivots), "root sequence differs from the oracle's"
        assert list(dp) == list(res.probs), "probabilities differ from the oracle's"
        assert ds["skipped"] == res.skipped and ds["nrounds"] == res.nrounds
        roots, probs = res.pivots[-128:], res.probs[-128:]
        e.reset_bc()
        applied, stats = e.batch(roots, probs, stats=True)
        assert applied == 128
        bc_batch = e.bc()
        print("[fullsize] batch done; oracle rounds", flush=True)
        for k in (0, 64, 127):
            lvl, sig, dl, _, ost = oracle.pivot(g, int(roots[k]), float(probs[k]), np.zeros(g.N))
            blvl, bsig = e.batch_levels(k)
            assert np.array_equal(blvl, lvl), f"root #{k} ({roots[k]}): lvl differs"
            assert np.array_equal(bsig, sig), f"root #{k} ({roots[k]}): sigma differs"
            ok, err = close(delta_of(e, k), dl)
            assert ok, f"root #{k} ({roots[k]}): delta rel err {err}"
            assert stats[k] == ost, f"root #{k} ({roots[k]}): counters {stats[k]} != {ost}"
            assert ost["nvisited"] > 1_000_000  # a full-size round
        print("[fullsize] plain rounds", flush=True)
        e.reset_bc()
        for v, p in zip(roots, probs):
            e.pivot(int(v), float(p))
        ok, err = close(bc_batch, e.bc())
        assert ok, f"batch vs plain bc rel err {err}"


def test_lattice_batch_against_oracle_and_plain(require_gpu):
    """The high-diameter case at full size: the 4096 x 4096 lattice with 3 %
    of its edges deleted (seed 3), ~8000 levels per round, 32 uniform roots
    in one batch of width 64.  Path counts there are binomial coefficients
    and wrap mod 2^32 (to 0 at times), so deltas hold infinities and NaNs:
    the comparison takes them at equal positions."""
    print("[fullsize] building the 4096^2 lattice", flush=True)
    g = Graph.lattice(4096, 0.03, seed=3, heuristic=0)
    L = _ffi.hip()
    with Engine(g) as e:
        assert L.axbc_set_batch_width(e.ctx, 64) == 0
        loop = PivotLoop(e, 32, c=0, x=1, z=0, batch=64)
        assert loop.step() == 32
        roots, probs = loop.res.pivots, loop.res.probs
        bc_batch = e.bc()
        print("[fullsize] lattice batch done; oracle rounds", flush=True)
        for k in (0, 31):
            lvl, sig, dl, _, ost = oracle.pivot(g, int(roots[k]), float(probs[k]), np.zeros(g.N))
            blvl, bsig = e.batch_levels(k)
            assert np.array_equal(blvl, lvl), f"root #{k} ({roots[k]}): lvl differs"
            assert np.array_equal(bsig, sig), f"root #{k} ({roots[k]}): sigma differs"
            ok, err = close(delta_of(e, k), dl)
            assert ok, f"root #{k} ({roots[k]}): delta rel err {err}"
            assert ost["nlevels"] > 4000  # a high-diameter round
        print("[fullsize] lattice plain rounds", flush=True)
        e.reset_bc()
        for v, p in zip(roots, probs):
            e.pivot(int(v), float(p))
        ok, err = close(bc_batch, e.bc())
        assert ok, f"batch vs plain bc rel err {err}"


def test_scale24_batch_against_oracle_and_plain(require_gpu):
    """The learned-proposal graph at full size: Kronecker SCALE 24 (16.8 M
    vertices, ~0.5 G directed entries), seed pair (4, 5), -H 1, 64 uniform
    roots in one batch of width 64: two roots' delta against the oracle and
    the batch's bc against one-root rounds on the device."""
    print("[fullsize] building SCALE 24 on the device", flush=True)
    g = Graph.kronecker_device(24, 16, seed=(4, 5), heuristic=1)
    L = _ffi.hip()
    with Engine(g) as e:
        assert L.axbc_set_batch_width(e.ctx, 64) == 0
        loop = PivotLoop(e, 512, c=0, x=1, z=0, batch=64)  # degree-0 draws consume attempts
        assert loop.step() == 64
        roots, probs = loop.res.pivots, loop.res.probs
        bc_batch = e.bc()
        print("[fullsize] SCALE 24 batch done; oracle rounds", flush=True)
        for k in (0, len(roots) - 1):
            lvl, sig, dl, _, _ = oracle.pivot(g, int(roots[k]), float(probs[k]), np.zeros(g.N))
            blvl, bsig = e.batch_levels(k)
            assert np.array_equal(blvl, lvl), f"root #{k} ({roots[k]}): lvl differs"
            assert np.array_equal(bsig, sig), f"root #{k} ({roots[k]}): sigma differs"
            ok, err = close(delta_of(e, k), dl)
            assert ok, f"root #{k} ({roots[k]}): delta rel err {err}"
        print("[fullsize] SCALE 24 plain rounds", flush=True)
        e.reset_bc()
        for v, p in zip(roots, probs):
            e.pivot(int(v), float(p))
        ok, err = close(bc_batch, e.bc())
        assert ok, f"batch vs plain bc rel err {err}"
