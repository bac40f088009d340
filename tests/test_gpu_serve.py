"""Serving: each job's result is the same alone and in a concurrent mix
(streams overlap on the device), admission respects the memory cap, and
oversize jobs are rejected."""
import numpy as np
import pytest

from axbc_amd.serve import Job, Server, footprint_bytes
from tests.parity import close

pytestmark = pytest.mark.gpu


def test_require(require_gpu):
    pass


def jobs():
    return [Job(jid=0, budget=60, z=6, scale=11, edgefactor=16, seed=(1, 2)),
            Job(jid=1, budget=80, z=1, scale=12, edgefactor=8, seed=(3, 4), batch=64, x=4, c=0),
            Job(jid=2, budget=50, z=0, scale=10, edgefactor=32, seed=(5, 6)),
            Job(jid=3, budget=70, z=4, scale=11, edgefactor=16, seed=(7, 8))]


def test_alone_equals_mix(require_gpu):
    print("[serve] alone", flush=True)
    alone = {}
    for j in jobs():
        srv = Server(slots=1)
        srv.submit(j)
        (d,) = srv.drain(timeout=300)
        srv.close()
        assert d.error is None, d.error
        alone[d.jid] = (d.bc, d.pivots)
    print("[serve] mix", flush=True)
    srv = Server(slots=4)
    for j in jobs():
        srv.submit(j)
    done = srv.drain(timeout=300)
    srv.close()
    assert len(done) == 4
    for d in done:
        assert d.error is None, d.error
        bc, piv = alone[d.jid]
        assert d.pivots == piv
        ok, err = close(d.bc, bc)
        assert ok, (d.jid, err)


def test_admission_by_footprint(require_gpu):
    js = jobs()[:3]
    for j in js:
        j.graph = None
    # cap that holds the largest job but not two of them at once
    from axbc_amd import Graph
    sizes = []
    for j in js:
        g = Graph.kronecker(j.scale, j.edgefactor, seed=j.seed, heuristic=j.heuristic)
        j.graph = g
        sizes.append(footprint_bytes(g.N, g.ned, j.batch))
    cap = max(sizes) + min(sizes) // 2
    srv = Server(slots=3, hbm_cap=cap)
    peak = [0]
    orig = srv._admit

    def watch():
        job = orig()
        peak[0] = max(peak[0], srv._used)
        return job

    srv._admit = watch
    for j in js:
        srv.submit(j)
    done = srv.drain(timeout=300)
    srv.close()
    assert all(d.error is None for d in done) and len(done) == 3
    assert peak[0] <= cap
    big = Job(jid=9, budget=10, z=0, scale=12, edgefactor=16, seed=(1, 2))
    srv = Server(slots=1, hbm_cap=1e5)
    srv.submit(big)
    (d,) = srv.drain(timeout=120)
    srv.close()
    assert d.error and "exceeds" in d.error


def test_concurrent_graph_builds(require_gpu):
    """Contexts on several threads rebuild their level-loop graphs (chunk
    changes) while the others launch, copy and build graphs on the device;
    every thread's batches must succeed and match its own single-thread
    result (graph construction touches no stream)."""
    import threading
    from axbc_amd import Engine, Graph, _ffi
    from axbc_amd.pivots import PivotLoop

    def work(seed, out):
        try:
            g = Graph.kronecker_device(12 + seed % 3, 16, seed=(seed + 1, seed + 2), heuristic=1)
            with Engine(g) as e:
                L = _ffi.hip()
                for chunk in (128, 256, 512, 1024, 128):
                    assert L.axbc_set_batch_chunk(e.ctx, chunk) == 0
                    loop = PivotLoop(e, 256, c=0, x=1, z=1, batch=64)
                    loop.run()
                out[seed] = (e.bc(), None)
        except Exception as ex:  # reported below
            out[seed] = (None, repr(ex))

    out = {}
    threads = [threading.Thread(target=work, args=(s, out)) for s in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    errors = {s: err for s, (_, err) in out.items() if err}
    assert not errors, errors
    alone = {}
    work(3, alone)
    ok, err = close(out[3][0], alone[3][0])
    assert ok, err


def test_serving_workload_vs_oracle(require_gpu):
    """The serving workload (Poisson arrivals; Kronecker SCALE 12-20, edge
    factor 8/16/32, budgets 64-2048, the -z 6 mixture in batches of 64
    roots) through four slots; every job with SCALE <= 16 matches the
    oracle's batch-synchronous pivot loop (sync_batch 64): same pivots,
    round count and bc within 1e-9."""
    from concurrent.futures import ThreadPoolExecutor

    from axbc_amd.serve import poisson_jobs
    from oracle import oracle

    js = [j for _, j in poisson_jobs(10, 50.0, seed=0)]
    assert {j.z for j in js} == {6} and {j.batch for j in js} == {64}
    assert min(j.scale for j in js) == 12 and max(j.scale for j in js) == 20
    srv = Server(slots=4, keep_bc=True)
    for j in js:
        srv.submit(j)
    done = {d.jid: d for d in srv.drain(timeout=600)}
    srv.close()
    assert len(done) == len(js)
    for d in done.values():
        assert d.error is None, (d.jid, d.error)
        assert d.nrounds > 0
    small = [d for d in done.values() if d.scale <= 16]
    assert len(small) >= 6
    print(f"[serve] {len(small)} jobs vs the oracle", flush=True)

    def ref(d):
        return oracle.run(d.graph, d.budget, c=d.c, x=d.x, z=d.z, sync_batch=d.batch)

    with ThreadPoolExecutor(max_workers=8) as ex:
        refs = list(ex.map(ref, small))
    for d, (pv, pr, bc, rs) in zip(small, refs):
        assert d.nrounds == rs["nrounds"], (d.jid, d.nrounds, rs)
        assert list(d.pivots) == [int(v) for v in pv], d.jid
        ok, err = close(d.bc, bc)
        assert ok, (d.jid, d.scale, err)
