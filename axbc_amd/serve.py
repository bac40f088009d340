"""Serving: BC-approximation jobs on device-resident graphs, admitted as they
arrive.

A job names a graph (Kronecker parameters or a prebuilt Graph), a pivot
budget and the pivot-loop flags.  The server keeps `slots` device workers;
each worker owns one engine context (its own HIP stream) at a time, so the
rounds of different jobs run concurrently on the GPU.  A job is admitted to
a free slot when its device footprint fits under the memory cap next to the
jobs already resident (jobs that do not fit wait; smaller jobs behind them
may pass).  Graph construction runs on host threads ahead of admission.
Latency of a job = completion time - arrival time.

No reference counterpart: AxBC runs one job per process.
"""
from __future__ import annotations

import ctypes as C
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _ffi
from .engine import BudgetError, Engine
from .graph import Graph
from .pivots import PivotLoop

PLAIN_STATE_BYTES = 4 + 4 + 8 + 8 + 4 + 4 + 4  # lvl, sigma, delta, bc, queue, qoff, reach per vertex
# device scratch of building a graph on the device (generator output, mirrored keys and three sort/select
# buffers of 8 B per mirrored edge, flags, radix-sort temporaries), per generated edge and per vertex
BUILD_BYTES_PER_EDGE = 112
BUILD_BYTES_PER_VERTEX = 48
SAMPLER_BYTES = 28 + 24 + 8                    # sampler scratch, LCC tables, slack per vertex


def build_bytes(scale: int, edgefactor: int) -> int:
    """Upper bound of the device scratch of Graph.kronecker_device, held
    while the graph is built (freed before the job is admitted)."""
    n = 1 << scale
    return n * edgefactor * BUILD_BYTES_PER_EDGE + n * BUILD_BYTES_PER_VERTEX


def footprint_bytes(N: int, ned: int, batch: int = 0, col: np.ndarray | None = None,
                    symmetric: bool = True) -> int:
    """Device bytes a job needs: CSC + per-vertex state + sampler tables,
    and when batched the batch state the library creates for this graph
    (axbc_batch_state_bytes: root rows, bit words, the level list at its
    first capacity, split-vertex and chunk rows, per-wave slots) plus the
    transpose of a graph that is not symmetric.  Without col the split
    vertices are estimated from a uniform degree."""
    b = 4 * (N + 1) + 4 * ned + N * (PLAIN_STATE_BYTES + SAMPLER_BYTES)
    if batch:
        width = 128 if batch > 64 else 64
        if col is None:
            col = np.minimum(np.arange(N + 1, dtype=np.uint64) * max(ned // max(N, 1), 1), ned).astype(np.uint32)
        size = C.c_size_t(0)
        rc = _ffi.hip().axbc_batch_state_bytes(N, ned, np.ascontiguousarray(col, np.uint32), width, C.byref(size))
        if rc != 0:
            raise ValueError(_ffi.hip().axbc_last_error().decode())
        b += int(size.value)
        if not symmetric:
            b += 4 * (N + 1 + ned)
    return b


@dataclass
class Job:
    jid: int
    budget: int                   # pivot attempts (-N)
    z: int = 6
    c: int = 0
    x: int = 1
    heuristic: int = 1
    batch: int = 0
    scale: int = 0
    edgefactor: int = 16
    seed: tuple = (1, 2)
    graph: Graph | None = None
    arrival: float = 0.0
    # results
    admitted: float = 0.0
    finished: float = 0.0
    nrounds: int = 0
    bc: np.ndarray | None = None
    pivots: list = field(default_factory=list)
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class Server:
    """slots device workers under a device memory cap of hbm_cap bytes.
    engine_factory(graph, device) makes a job's engine (an Engine),
    runner(engine, job) runs its pivot loop and builder(job, device) builds a
    job's graph on the device; tests pass stand-ins to exercise admission
    without a GPU.  A device build holds its scratch (build_bytes) under the
    same cap while it runs."""

    def __init__(self, slots: int = 4, hbm_cap: float = 200e9, device: int = 0, build_threads: int = 4,
                 keep_bc: bool = True, engine_factory=Engine, runner=None, builder=None):
        self.slots, self.cap, self.device, self.keep_bc = slots, hbm_cap, device, keep_bc
        self._factory = engine_factory
        self._runner = runner or self._pivot_loop
        self._builder_fn = builder or self._kronecker
        self._ready: list[Job] = []      # built, waiting for admission (arrival order)
        self._lock = threading.Condition()
        self._used = 0
        self._active = 0
        self._done: list[Job] = []
        self._pending = 0
        self._stop = False
        self._builder = ThreadPoolExecutor(max_workers=build_threads)
        self._workers = [threading.Thread(target=self._work, daemon=True) for _ in range(slots)]
        for w in self._workers:
            w.start()

    # -- client side ------------------------------------------------------
    def submit(self, job: Job) -> None:
        job.arrival = job.arrival or time.perf_counter()
        with self._lock:
            self._pending += 1
        self._builder.submit(self._build, job)

    def drain(self, timeout: float | None = None) -> list[Job]:
        t0 = time.perf_counter()
        with self._lock:
            while self._pending:
                left = None if timeout is None else timeout - (time.perf_counter() - t0)
                if left is not None and left <= 0:
                    break
                self._lock.wait(left if left is not None else 1.0)
            return list(self._done)

    def close(self) -> None:
        with self._lock:
            self._stop = True
            self._lock.notify_all()
        for w in self._workers:
            w.join()
        self._builder.shutdown(wait=True)

    # -- internals ----------------------------------------------------------
    def _kronecker(self, job: Job, device: int) -> Graph:
        return Graph.kronecker_device(job.scale, job.edgefactor, seed=job.seed, heuristic=job.heuristic,
                                      device=device)

    def _build(self, job: Job) -> None:
        scratch = build_bytes(job.scale, job.edgefactor) if job.graph is None else 0
        if scratch > self.cap:
            job.error = f"graph build needs {scratch} B of device scratch, which exceeds the memory cap {int(self.cap)} B"
            scratch = 0
        elif scratch:
            with self._lock:  # the build's scratch counts against the cap while it runs
                while self._used + scratch > self.cap and not self._stop:
                    self._lock.wait(0.05)
                self._used += scratch
        try:
            if job.graph is None and not job.error:  # generated and preprocessed on the device
                job.graph = self._builder_fn(job, self.device)
        except Exception as ex:  # the job fails, the server goes on
            job.error = f"graph build failed: {type(ex).__name__}: {ex}"
        finally:
            if scratch:
                with self._lock:
                    self._used -= scratch
                    self._lock.notify_all()
        with self._lock:
            self._ready.append(job)
            self._lock.notify_all()

    def _admit(self) -> Job | None:
        """First built job (arrival order) whose footprint fits; a job larger
        than the whole cap is failed immediately."""
        for i, job in enumerate(self._ready):
            if job.error:
                return self._ready.pop(i)
            need = getattr(job, "_retry_need", 0) or footprint_bytes(
                job.graph.N, job.graph.ned, job.batch, getattr(job.graph, "col", None))
            if need > self.cap:
                job.error = f"footprint {need} B exceeds the memory cap {int(self.cap)} B"
                return self._ready.pop(i)
            if self._used + need <= self.cap:
                job._need = need
                self._used += need
                return self._ready.pop(i)
        return None

    def _work(self) -> None:
        while True:
            with self._lock:
                job = None
                while job is None:
                    if self._stop:
                        return
                    job = self._admit()
                    if job is None:
                        self._lock.wait(0.05)
                self._active += 1
            job.admitted = job.admitted or time.perf_counter()
            retry = False
            try:
                if not job.error:
                    self._run(job)
            except BudgetError as ex:
                # the reservation was too small: admit once more with twice it
                if getattr(job, "_retry_need", 0):
                    job.error = f"BudgetError: {ex}"
                else:
                    job._retry_need = 2 * job._need
                    retry = True
            except Exception as ex:
                job.error = f"{type(ex).__name__}: {ex}"
            job.finished = time.perf_counter()
            with self._lock:
                self._used -= getattr(job, "_need", 0)
                job._need = 0
                self._active -= 1
                if retry:
                    self._ready.insert(0, job)
                else:
                    self._pending -= 1
                    self._done.append(job)
                self._lock.notify_all()

    def _charge(self, job: Job, actual: int) -> None:
        """The context exists: charge its actual device bytes where they
        exceed the reservation."""
        with self._lock:
            if actual > job._need:
                self._used += actual - job._need
                job._need = actual

    def _run(self, job: Job) -> None:
        with self._factory(job.graph, self.device) as e:
            self._charge(job, e.device_bytes)
            e.set_budget(job._need)
            self._runner(e, job)
        if not self.keep_bc:
            job.graph = None

    def _pivot_loop(self, e: Engine, job: Job) -> None:
        loop = PivotLoop(e, job.budget, c=job.c, x=job.x, z=job.z, batch=job.batch)
        res = loop.run()
        job.nrounds = res.nrounds
        job.pivots = res.pivots
        if self.keep_bc:
            job.bc = e.bc()


def poisson_jobs(n: int, rate: float, seed: int = 0, scales=(12, 20), efs=(8, 16, 32), budget=(64, 2048),
                 z: int = 6, batch: int = 64) -> list[tuple[float, Job]]:
    """The serving workload: Poisson arrivals at `rate` per second; per job a
    Kronecker SCALE drawn uniformly from `scales`, an edge factor from `efs`,
    the job id as seed, and a pivot budget log-uniform in `budget`.  Jobs
    run in bit-packed batches of `batch` roots (batch-synchronous for the
    dynamic distributions; 0 = one root per round)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for j in range(n):
        t += rng.exponential(1.0 / rate)
        s = int(rng.integers(scales[0], scales[1] + 1))
        ef = int(rng.choice(efs))
        b = int(round(np.exp(rng.uniform(np.log(budget[0]), np.log(budget[1])))))
        out.append((t, Job(jid=j, budget=b, z=z, scale=s, edgefactor=ef, seed=(j + 1, j + 2), batch=batch)))
    return out


def run_workload(jobs: list[tuple[float, Job]], slots: int = 4, hbm_cap: float = 200e9,
                 keep_bc: bool = False) -> dict:
    """Replay arrivals in real time and report latency percentiles and the
    aggregate pivot throughput."""
    srv = Server(slots=slots, hbm_cap=hbm_cap, keep_bc=keep_bc)
    t0 = time.perf_counter()
    for at, job in jobs:
        now = time.perf_counter() - t0
        if at > now:
            time.sleep(at - now)
        job.arrival = time.perf_counter()
        srv.submit(job)
    done = srv.drain()
    t1 = time.perf_counter()
    srv.close()
    ok = [j for j in done if not j.error]
    lat = np.array([j.latency for j in ok]) if ok else np.zeros(1)
    rounds = sum(j.nrounds for j in ok)
    errors: dict[str, int] = {}
    for j in done:
        if j.error:
            errors[j.error] = errors.get(j.error, 0) + 1
    return {
        "jobs": len(done), "failed": len(done) - len(ok), "errors": errors,
        "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
        "mean_s": float(lat.mean()), "pivots": rounds, "wall_s": t1 - t0,
        "pivots_per_s": rounds / (t1 - t0) if t1 > t0 else 0.0,
    }
