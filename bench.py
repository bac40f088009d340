"""Benchmark of the sampled-Brandes hot path on one MI355X.

Workload (the batched configuration of the reference's headline metric):
Kronecker SCALE 22, edge factor 16, A/B/C/D = .57/.19/.19/.05, generator
seed pair (2, 3), symmetrised, 1-degree reduction (-H 1), adaptive stop
(-c 1 -x 1024), high-degree-neighbourhood pivots (-z 1), at most 4096
pivots.  A step is one batch of B roots of the pivot loop (sampling, stop
test, forward + backward sweep, bc update); the plain one-root-per-round
path runs the same loop for a fixed number of rounds for the batched/plain
ratio.  The CPU oracle (plain C, one thread) is timed on a bounded sample
of the same graph and pivots.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

WORKLOAD = {
    "workload": "kronecker-s22-ef16-H1-z1-c1-x1024",
    "graph": "Kronecker SCALE 22 (4,194,304 vertices), edgefactor 16, A/B/C/D=.57/.19/.19/.05, seed (2,3), "
             "symmetrized, 1-degree reduction",
    "pivots": "high-degree-neighbourhood sampling (-z 1), srand48(4343543), adaptive stop c=1 x=1024, max 4096",
}
KIND_NAMES = {0: "k_bpush", 1: "k_bsigma", 2: "k_bclose", 3: "k_bdep", 4: "k_bupdate", 5: "k_fwd_expand",
              6: "k_bwd_pull", 7: "k_copy", 8: "k_bpush_split", 9: "k_bsigma_split", 10: "k_bdep_split",
              11: "k_bdep_combine", 12: "k_fwd_hubs", 13: "k_bwd_hubs", 14: "k_bpull", 15: "k_bpull_split"}
BATCH_KINDS = (0, 8, 1, 9, 14, 15, 2, 3, 10, 11, 4)
PLAIN_KINDS = (5, 12, 6, 13)
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (KiB per dispatch) on this workload, scaled by the same counters' reading
# for a copy kernel of known size.
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_stats(e, kind):
    from axbc_amd import _ffi
    L = _ffi.hip()
    ms, n, b = C.c_double(), C.c_uint64(), C.c_uint64()
    L.axbc_kernel_stats.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64)]
    rc = L.axbc_kernel_stats(e.ctx, kind, C.byref(ms), C.byref(n), C.byref(b))
    if rc != 0:
        raise RuntimeError(L.axbc_last_error())
    return ms.value, n.value, b.value


def set_timing(e, on):
    from axbc_amd import _ffi
    L = _ffi.hip()
    L.axbc_kernel_timing.argtypes = [C.c_void_p, C.c_int]
    if L.axbc_kernel_timing(e.ctx, 1 if on else 0) != 0:
        raise RuntimeError(L.axbc_last_error())


def log_kernels(kstats):
    for k, (ms, n, b) in kstats.items():
        if n:
            log(f"  {KIND_NAMES[k]}: {n} launches, {ms:.2f} ms, avg {ms / n * 1e3:.1f} us, "
                f"{b / 1e9:.3f} GB -> {b / (ms * 1e-3) / 1e9 if ms else 0:.1f} GB/s")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary,
    or None when the kernel was not measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k or t.get("workload") != WORKLOAD["workload"]:
        return None
    return int(k["hbm_bytes_per_launch"])


def copy_calibration(mib):
    """The calibration copy kernel (reads and writes `mib` MiB per launch),
    timed with HIP events; used to scale the hardware byte counters."""
    from axbc_amd import Engine, Graph, _ffi
    g = Graph.kronecker(10, 16, seed=(1, 2))
    e = Engine(g)
    L = _ffi.hip()
    set_timing(e, True)
    if L.axbc_copy_kernel(e.ctx, mib << 20, 8) != 0:
        raise RuntimeError(L.axbc_last_error())
    ms, n, b = kernel_stats(e, 7)
    set_timing(e, False)
    e.close()
    print(json.dumps({"kernel": "k_copy", "bytes_read_per_launch": mib << 20, "bytes_written_per_launch": mib << 20,
                      "launches": n, "avg_launch_us": round(ms / n * 1e3, 2),
                      "gb_per_s": round(b / (ms * 1e-3) / 1e9, 1)}))


def cpu_baseline(g, roots, probs, seconds):
    """The oracle (one thread) on the same graph and root sequence, bounded
    to about `seconds` of work."""
    from oracle import oracle
    bc = np.zeros(g.N)
    done, er = 0, 0
    t0 = time.perf_counter()
    for v, p in zip(roots, probs):
        _, _, _, bc, st = oracle.pivot(g, int(v), float(p), bc)
        done += 1
        er += st["e_r"]
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round(done / dt, 4), "unit": "pivots/s", "cores": 1, "kind": "port",
            "sample": f"{done} pivots of the same workload (first roots of the loop), oracle/brandes_oracle.c "
                      f"single-threaded, {dt:.1f} s; GTEPS {2 * er / dt / 1e9:.4f}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24, help="timed batches")
    ap.add_argument("--warmup", type=int, default=2, help="untimed batches")
    ap.add_argument("--batch", type=int, default=0, choices=[0, 32, 64, 128],
                    help="roots per batch; 0 = chosen by the autotuner (with the split and pull settings)")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--plain-rounds", type=int, default=48)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every level kernel from the host (the timed batches always do, for the per-kernel "
                         "timers; this also covers the autotuning trials)")
    ap.add_argument("--chunk", type=int, default=0, help="split threshold with --batch (default 256)")
    ap.add_argument("--pull", type=int, default=-1, help="bottom-up ratio x16 with --batch (default 32)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the calibration copy kernel over MIB MiB (8 launches) and print its rate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs (pivot sharding with one "
                                   "all-reduce of bc) are not part of this build", "n_gpus": args.gpus}))
        sys.exit(2)

    from axbc_amd import Engine, Graph
    from axbc_amd.pivots import PivotLoop

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    t0 = time.perf_counter()
    g = Graph.kronecker(args.scale, 16, seed=(2, 3), heuristic=1)
    log(f"graph: N={g.N} ned={g.ned} built in {time.perf_counter() - t0:.1f}s")
    e = Engine(g)
    log(f"device bytes {e.device_bytes >> 20} MB")
    if args.no_graphs:
        from axbc_amd import _ffi
        if _ffi.hip().axbc_set_batch_graphs(e.ctx, 0) != 0:
            raise RuntimeError(_ffi.hip().axbc_last_error())

    # ---- autotuning (untimed): batch width, split threshold, pull ratio ----
    tuned = None
    batch = args.batch
    if not batch:
        from axbc_amd.autotune import tune
        rng = np.random.default_rng(0)
        sample = rng.choice(np.flatnonzero(np.diff(g.col) > 0), 128 * 3, replace=False)
        t = time.perf_counter()
        tuned = tune(e, sample, reps=1)
        batch = tuned["width"]
        log(f"autotune {time.perf_counter() - t:.1f}s: width {tuned['width']} chunk {tuned['chunk']} "
            f"pull {tuned['pull']} ({tuned['s_per_root'] * 1e3:.3f} ms/root)")
        e.reset_bc()
    else:
        from axbc_amd import _ffi
        L = _ffi.hip()
        if args.chunk and L.axbc_set_batch_chunk(e.ctx, args.chunk) != 0:
            raise RuntimeError(L.axbc_last_error())
        if args.pull >= 0 and L.axbc_set_batch_pull(e.ctx, args.pull) != 0:
            raise RuntimeError(L.axbc_last_error())

    # ---- batched pivot loop ---------------------------------------------
    loop = PivotLoop(e, 4096, c=1, x=1024, z=1, batch=batch)
    for _ in range(args.warmup):
        loop.step()
    e.sync()
    set_timing(e, True)
    rounds0, er0 = loop.res.nrounds, loop.res.e_r
    steps = 0
    t = time.perf_counter()
    for _ in range(args.steps):
        if not loop.step():
            break
        steps += 1
    e.sync()
    dt = time.perf_counter() - t
    rounds = loop.res.nrounds - rounds0
    er = loop.res.e_r - er0
    kstats = {k: kernel_stats(e, k) for k in BATCH_KINDS}
    set_timing(e, False)
    pivots_s = rounds / dt
    gteps = 2.0 * er / dt / 1e9
    log(f"batched: {rounds} rounds in {dt:.3f}s -> {pivots_s:.1f} pivots/s, {gteps:.2f} GTEPS")
    log_kernels(kstats)
    dom = max(kstats, key=lambda k: kstats[k][0])
    ms, n, b = kstats[dom]
    achieved = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(KIND_NAMES[dom]),
                "kernel": KIND_NAMES[dom], "avg_launch_us": round(ms / max(n, 1) * 1e3, 2), "launches": n,
                "algorithmic_bytes_per_launch": int(b / max(n, 1)),
                "kernel_ms_share": round(ms / max(sum(v[0] for v in kstats.values()), 1e-9), 3)}

    # ---- plain one-root-per-round loop on the same engine --------------------
    e.reset_bc()
    ploop = PivotLoop(e, 4096, c=0, x=1024, z=1, batch=0)
    for _ in range(4):
        ploop.step()
    e.sync()
    r0p, er0p = ploop.res.nrounds, ploop.res.e_r
    set_timing(e, True)
    t = time.perf_counter()
    for _ in range(args.plain_rounds):
        if not ploop.step():
            break
    e.sync()
    dtp = time.perf_counter() - t
    log_kernels({k: kernel_stats(e, k) for k in PLAIN_KINDS})
    set_timing(e, False)
    prounds = ploop.res.nrounds - r0p
    plain_s = prounds / dtp
    plain_gteps = 2.0 * (ploop.res.e_r - er0p) / dtp / 1e9
    log(f"plain: {prounds} rounds in {dtp:.3f}s -> {plain_s:.2f} pivots/s, {plain_gteps:.2f} GTEPS")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(g, ploop.res.pivots, ploop.res.probs, args.cpu_seconds)
        log(f"cpu: {cpu}")
    e.close()

    out = {
        "metric": "pivots/s & GTEPS (fwd+bwd) on 1 MI355X, batched vs plain; p99 request latency",
        "value": round(pivots_s, 3),
        "unit": "pivots/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / max(steps, 1) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 sigma / f64 delta,bc",
        "data": "synthetic (Kronecker generator identical to the reference's Graph500 generator)",
        "config": dict(WORKLOAD, batch_width=batch, pivots_per_step=batch,
                       autotune={k: tuned[k] for k in ("width", "chunk", "pull")} if tuned else None),
        "gteps": round(gteps, 3),
        "plain_pivots_per_s": round(plain_s, 3),
        "plain_gteps": round(plain_gteps, 3),
        "batched_over_plain": round(pivots_s / plain_s, 3) if plain_s else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
