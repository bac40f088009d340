"""The context's device-memory budget (axbc_set_memory_budget): an
allocation that would pass it fails with AXBC_EBUDGET before anything is
allocated, bc stays untouched, the bytes in use stay what they were, and
the context runs normally once the budget allows it."""
import numpy as np
import pytest

from axbc_amd import Engine, Graph
from axbc_amd.engine import BudgetError
from oracle import oracle
from tests.parity import close

pytestmark = pytest.mark.gpu


def test_require(require_gpu):
    pass


def path_graph(n):
    e = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1)
    return Graph.from_edges(n, e, undirected=True, heuristic=0, name=f"path{n}")


def bc_of_roots(g, roots):
    bc = np.zeros(g.N)
    for v in roots:
        bc = oracle.pivot(g, int(v), 1.0 / g.N, bc)[3]
    return bc


def test_batch_state_over_budget(require_gpu):
    g = Graph.kronecker(12, 16, seed=(1, 2), heuristic=1)
    roots = np.arange(64, dtype=np.uint32) * 7 % g.N
    probs = np.full(64, 1.0 / g.N)
    with Engine(g) as e:
        base = e.bytes_in_use()
        assert base == e.device_bytes > 0
        e.set_budget(base + 4096)
        with pytest.raises(BudgetError, match="the batch state"):
            e.batch(roots, probs)
        assert e.bytes_in_use() == base
        assert not e.bc().any()
        e.set_budget(0)
        applied, _ = e.batch(roots, probs)
        assert applied == 64
        assert e.bytes_in_use() > base
        ok, err = close(e.bc(), bc_of_roots(g, roots))
        assert ok, err


def test_level_list_growth_over_budget(require_gpu):
    """64 roots spread along a path list most vertices at 64 distinct
    levels: the level list outgrows its first 8N entries.  With the budget
    at the bytes of the first batch state, the growth fails cleanly."""
    g = path_graph(4000)
    roots = (np.arange(64, dtype=np.uint32) * 61 + 5) % g.N
    probs = np.full(64, 1.0 / g.N)
    with Engine(g) as e:
        e.set_batch_mode(1)  # the bit-packed sweep (its level list)
        e.batch(roots[:1], probs[:1])  # creates the batch state at its first capacity
        first = e.bytes_in_use()
        e.reset_bc()
        e.set_budget(first)
        with pytest.raises(BudgetError, match="the batch state"):
            e.batch(roots, probs)
        assert not e.bc().any()
        assert e.bytes_in_use() <= first
        e.set_budget(0)
        applied, _ = e.batch(roots, probs)
        assert applied == 64
        assert e.bytes_in_use() > first
        ok, err = close(e.bc(), bc_of_roots(g, roots))
        assert ok, err


def test_workgroup_sweep_state_over_budget(require_gpu):
    """The workgroup sweep allocates its per-root state on first use: with
    the budget at the bytes in use after a bit-packed batch, the first
    workgroup-sweep batch fails cleanly (bc untouched); with no budget it
    runs and matches the oracle."""
    g = path_graph(3000)
    roots = (np.arange(64, dtype=np.uint32) * 37 + 3) % g.N
    probs = np.full(64, 1.0 / g.N)
    with Engine(g) as e:
        e.set_batch_mode(1)
        e.batch(roots[:8], probs[:8])
        e.reset_bc()
        first = e.bytes_in_use()
        e.set_budget(first)
        e.set_batch_mode(2)
        with pytest.raises(BudgetError, match="workgroup sweep"):
            e.batch(roots, probs)
        assert not e.bc().any()
        assert e.bytes_in_use() <= first
        e.set_budget(0)
        applied, _ = e.batch(roots, probs)
        assert applied == 64
        assert e.bytes_in_use() > first
        ok, err = close(e.bc(), bc_of_roots(g, roots))
        assert ok, err


def test_budget_below_use_refuses_only_new_allocations(require_gpu):
    g = Graph.kronecker(11, 16, seed=(3, 4), heuristic=1)
    roots = np.arange(32, dtype=np.uint32) * 5 % g.N
    probs = np.full(32, 1.0 / g.N)
    with Engine(g) as e:
        e.batch(roots, probs)
        held = e.bytes_in_use()
        e.set_budget(1)  # below what is held: nothing is freed, nothing more allocated
        applied, _ = e.batch(roots, probs)  # state exists: no allocation
        assert applied == 32
        with pytest.raises(BudgetError):
            e.batch_delta(0)  # the per-root row copy would be a new allocation
        assert e.bytes_in_use() == held
        e.set_budget(0)
        e.batch_delta(0)


@pytest.mark.parametrize("width", [64, 128, 256])
def test_batch_state_bytes_are_exact(require_gpu, width):
    """The admission estimate of the batch state equals what the first
    batch allocates (at width 256, a batch of more than 128 roots: both
    states)."""
    import ctypes as C

    from axbc_amd import _ffi
    g = Graph.kronecker(12, 16, seed=(5, 6), heuristic=1)
    size = C.c_size_t(0)
    assert _ffi.hip().axbc_batch_state_bytes(g.N, g.ned, g.col, width, C.byref(size)) == 0
    with Engine(g) as e:
        e.set_batch_width(width)
        base = e.bytes_in_use()
        nb = 200 if width == 256 else 4
        e.set_batch_mode(1)
        e.batch(np.arange(nb, dtype=np.uint32), np.full(nb, 1.0 / g.N))
        assert e.bytes_in_use() - base == size.value


def test_second_state_over_budget(require_gpu):
    """A batch of more than 128 bit-packed roots needs a second batch state:
    with the budget at the bytes held after a batch of 128, the wider batch
    fails cleanly (bc untouched, nothing more held), and a batch of 128
    still runs."""
    g = Graph.kronecker(12, 16, seed=(7, 9), heuristic=1)
    deg = g.degrees()
    roots = np.random.default_rng(4).choice(np.flatnonzero(deg > 0), size=256, replace=False).astype(np.uint32)
    probs = np.full(256, 1.0 / g.N)
    with Engine(g) as e:
        e.set_batch_mode(1)
        e.set_batch_width(256)
        e.batch(roots[:128], probs[:128])  # the first state only
        held = e.bytes_in_use()
        e.reset_bc()
        e.set_budget(held)
        with pytest.raises(BudgetError, match="the batch state"):
            e.batch(roots, probs)
        assert not e.bc().any()
        assert e.bytes_in_use() == held
        applied, _ = e.batch(roots[:128], probs[:128])
        assert applied == 128
        e.set_budget(0)
        applied, _ = e.batch(roots, probs)
        assert applied == 256
        assert e.bytes_in_use() > held
        bc_ref = bc_of_roots(g, np.concatenate([roots[:128], roots]))
        ok, err = close(e.bc(), bc_ref)
        assert ok, err
