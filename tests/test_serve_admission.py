"""Serving admission and the device-memory budget, on the CPU with a
stand-in engine: jobs are admitted by footprint under the cap, a job larger
than the cap fails at once, the server charges a context's actual bytes,
and a job whose allocations pass its reservation (AXBC_EBUDGET) is admitted
once more with twice the reservation before it fails."""
import threading
import time
from dataclasses import dataclass

import numpy as np

from axbc_amd.engine import BudgetError
from axbc_amd.serve import Job, Server, footprint_bytes


@dataclass
class FakeGraph:
    N: int
    ned: int


class Ledger:
    """Device bytes held by the live stand-in engines."""

    def __init__(self):
        self.lock = threading.Lock()
        self.live = 0
        self.peak = 0

    def add(self, n):
        with self.lock:
            self.live += n
            self.peak = max(self.peak, self.live)

    def sub(self, n):
        with self.lock:
            self.live -= n


class FakeEngine:
    """Holds device_bytes from creation (extra_create more than the estimate
    for graphs tagged so) and allocates more through grow(), which enforces
    the budget as the engine does."""

    def __init__(self, ledger, g, device=0):
        self.ledger, self.g = ledger, g
        self.device_bytes = 4 * (g.N + 1) + 4 * g.ned + getattr(g, "extra_create", 0)
        self.held = self.device_bytes
        self.budget = 0
        ledger.add(self.held)

    def set_budget(self, n):
        self.budget = n

    def grow(self, n):
        if self.budget and self.held + n > self.budget:
            raise BudgetError(f"axbc error -5: device memory budget: {n} more, {self.held} of {self.budget}")
        self.held += n
        self.ledger.add(n)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.ledger.sub(self.held)


def runner(e, job):
    # the job's working set after creation: graph.work bytes, then some time
    e.grow(getattr(job.graph, "work", 0))
    time.sleep(0.01)
    job.nrounds = job.budget


def serve(jobs, cap, slots=4):
    ledger = Ledger()
    srv = Server(slots=slots, hbm_cap=cap, build_threads=1, keep_bc=False,
                 engine_factory=lambda g, d: FakeEngine(ledger, g, d), runner=runner)
    for j in jobs:
        srv.submit(j)
    done = srv.drain(timeout=30)
    srv.close()
    return {j.jid: j for j in done}, ledger


def job(jid, N, ned, work=0, budget=10, extra_create=0):
    g = FakeGraph(N, ned)
    g.work = work
    g.extra_create = extra_create
    return Job(jid=jid, budget=budget, graph=g)


def test_admission_respects_the_cap():
    N, ned = 1000, 16000
    need = footprint_bytes(N, ned)
    cap = int(2.5 * need)  # at most two such jobs at once
    done, ledger = serve([job(i, N, ned, work=need // 2) for i in range(12)], cap, slots=6)
    assert len(done) == 12 and all(not j.error for j in done.values())
    assert ledger.peak <= cap
    assert ledger.live == 0


def test_job_larger_than_the_cap_fails_at_once():
    small = footprint_bytes(100, 1000)
    done, _ = serve([job(0, 100, 1000), job(1, 10 ** 6, 10 ** 7), job(2, 100, 1000)], cap=3 * small)
    assert not done[0].error and not done[2].error
    assert "exceeds the memory cap" in done[1].error


def test_allocation_past_the_reservation_is_readmitted_once():
    N, ned = 1000, 16000
    need = footprint_bytes(N, ned)
    create = 4 * (N + 1) + 4 * ned
    cap = 10 * need
    jobs = [job(0, N, ned, work=need - create + 1),        # just past the reservation: second try fits
            job(1, N, ned, work=2 * need - create + 1),    # past twice the reservation: fails
            job(2, N, ned, work=need // 4)]                # fits at once
    done, ledger = serve(jobs, cap)
    assert not done[0].error and done[0].nrounds == 10
    assert done[1].error and done[1].error.startswith("BudgetError")
    assert not done[2].error
    assert ledger.live == 0


def test_actual_bytes_above_the_estimate_are_charged():
    # contexts that hold more at creation than the estimate: the server
    # charges the actual bytes, so no more run together than fit
    N, ned = 1000, 16000
    need = footprint_bytes(N, ned)
    cap = 3 * need
    jobs = [job(i, N, ned, extra_create=need) for i in range(6)]  # each holds ~2x the estimate
    done, ledger = serve(jobs, cap, slots=6)
    assert all(not j.error for j in done.values())
    # two jobs of ~2x need would hold ~4x need > cap; charging keeps the
    # sum of reservations under the cap, but the first job's excess is only
    # known after its creation, so the peak is at most the cap plus one excess
    assert ledger.peak <= cap + need + 4 * (N + 1) + 4 * ned


def test_poisson_jobs_distribution():
    from axbc_amd.serve import poisson_jobs
    js = poisson_jobs(400, 20.0, seed=1)
    t = np.array([a for a, _ in js])
    assert np.all(np.diff(t) > 0) and 15 < 400 / t[-1] < 25
    sc = np.array([j.scale for _, j in js])
    assert sc.min() == 12 and sc.max() == 20
    b = np.array([j.budget for _, j in js])
    assert b.min() >= 64 and b.max() <= 2048
    assert {j.edgefactor for _, j in js} == {8, 16, 32}


def test_device_builds_hold_their_scratch_under_the_cap():
    """Graph builds on the device reserve build_bytes while they run: with
    a cap for one build at a time, builds never overlap, and a build larger
    than the cap fails at once."""
    from axbc_amd.serve import build_bytes
    ledger = Ledger()
    live = []
    peak = [0]
    lock = threading.Lock()

    def builder(job, device):
        with lock:
            live.append(job.jid)
            peak[0] = max(peak[0], len(live))
        time.sleep(0.02)
        with lock:
            live.remove(job.jid)
        return FakeGraph(1 << job.scale, 16 << job.scale)

    cap = int(1.5 * build_bytes(12, 16))
    srv = Server(slots=2, hbm_cap=cap, build_threads=4, keep_bc=False,
                 engine_factory=lambda g, d: FakeEngine(ledger, g, d), runner=runner, builder=builder)
    for i in range(6):
        srv.submit(Job(jid=i, budget=5, scale=12, edgefactor=16))
    srv.submit(Job(jid=99, budget=5, scale=24, edgefactor=32))
    done = {d.jid: d for d in srv.drain(timeout=30)}
    srv.close()
    assert all(not done[i].error for i in range(6)), [done[i].error for i in range(6)]
    assert peak[0] == 1
    assert "graph build" in done[99].error and "exceeds the memory cap" in done[99].error
