"""Benchmark of the sampled-Brandes hot path on one MI355X.

Workload (the batched configuration of the reference's headline metric):
Kronecker SCALE 22, edge factor 16, A/B/C/D = .57/.19/.19/.05, generator
seed pair (2, 3), symmetrised, 1-degree reduction (-H 1), adaptive stop
(-c 1 -x 1024), high-degree-neighbourhood pivots (-z 1), at most 4096
pivots.  A step is one batch of B roots of the pivot loop (sampling, stop
test, forward + backward sweep, bc update); the plain one-root-per-round
path runs the same loop for a fixed number of rounds for the batched/plain
ratio.  The CPU oracle (plain C) is timed on a bounded sample of the same
graph and pivots, on one thread and over the host's OpenMP threads.

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
KIND_NAMES = {0: "k_bpush", 1: "k_bsigma", 2: "k_bclose", 3: "k_bdep", 4: "k_bupdate", 5: "k_fwd_level",
              6: "k_bwd_pull", 7: "k_copy", 8: "k_bpush_split", 9: "k_bsigma_split", 10: "k_bdep_split",
              11: "k_bdep_combine", 12: "k_fwd_hub_level", 13: "k_bwd_hubs", 14: "k_bpull", 15: "k_bpull_split",
              16: "k_wgsweep", 17: "k_fwd_pull", 18: "k_fwd_pull_hubs", 19: "k_read", 20: "k_gather"}
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


def set_timing_kinds(e, kinds):
    """Events and byte counters on the given kernel kinds only."""
    from axbc_amd import _ffi
    L = _ffi.hip()
    mask = 0
    for k in kinds:
        mask |= 1 << k
    if L.axbc_kernel_timing_kinds(e.ctx, mask) != 0:
        raise RuntimeError(L.axbc_last_error())


def log_kernels(kstats):
    """One line per kernel kind; the plain kernels count no bytes."""
    for k, (ms, n, b) in kstats.items():
        if n:
            rate = f", {b / 1e9:.3f} GB -> {b / (ms * 1e-3) / 1e9 if ms else 0:.1f} GB/s" if b else ""
            log(f"  {KIND_NAMES[k]}: {n} launches, {ms:.2f} ms, avg {ms / n * 1e3:.1f} us{rate}")


def measured_traffic(kernel, batch, chunk, pull):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (one entry per batch width and split threshold the autotuner can pick),
    or None when the kernel was not measured at these settings."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != WORKLOAD["workload"]:
        return None
    # the backward sweep and the update do not depend on the forward direction switch
    backward = kernel.startswith("k_bdep") or kernel == "k_bupdate"
    for s in t.get("settings", []):
        if (s.get("batch_width"), s.get("chunk")) != (batch, chunk) or (not backward and s.get("pull") != pull):
            continue
        k = s.get("kernels", {}).get(kernel)
        return int(k["hbm_bytes_per_launch"]) if k else None
    return None


def roofline_alone(e, dom, no_graphs):
    """The dominant kernel's rate on batches of 128 roots (one sweep, one
    stream) of the workload's distribution: 2 warm-up and 6 timed batches."""
    from axbc_amd import _ffi
    from axbc_amd.pivots import PivotLoop
    L = _ffi.hip()
    roots = PivotLoop(e, 4096, c=0, x=1024, z=1, seed=4343543 + 2).sample_roots(128 * 8)
    saved = e.bc()
    if L.axbc_set_batch_width(e.ctx, 128) != 0:
        raise RuntimeError(L.axbc_last_error())
    p = np.full(128, 1.0 / e.N)
    for i in range(2):
        e.batch(roots[128 * i:128 * (i + 1)], p)
    e.sync()
    set_timing_kinds(e, [dom])
    for i in range(2, 8):
        e.batch(roots[128 * i:128 * (i + 1)], p)
    e.sync()
    ms, n, b = kernel_stats(e, dom)
    set_timing(e, False)
    e.set_bc(saved)
    achieved = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"batch_width": 128, "achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
            "avg_launch_us": round(ms / max(n, 1) * 1e3, 2), "launches": n,
            "algorithmic_bytes_per_launch": int(b / max(n, 1))}


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
    if L.axbc_read_kernel(e.ctx, mib << 20, 8) != 0:
        raise RuntimeError(L.axbc_last_error())
    rms, rn, rb = kernel_stats(e, 19)
    # random 128-byte line fetches over regions of growing size: one XCD's
    # L2 (4 MiB), the eight L2s, the 256 MiB Infinity Cache, and the 4 GiB
    # of a SCALE 22 dependency row array at width 128
    gather = []
    for wmib in (4, 32, 256, 4096):
        set_timing(e, False)
        set_timing(e, True)
        if L.axbc_gather_kernel(e.ctx, wmib << 20, 8) != 0:
            raise RuntimeError(L.axbc_last_error())
        gms, gn, gb = kernel_stats(e, 20)
        gather.append({"window_mib": wmib, "launches": gn, "avg_launch_us": round(gms / gn * 1e3, 2),
                       "glines_per_s": round(gb / 128 / (gms * 1e-3) / 1e9, 2),
                       "gb_per_s": round(gb / (gms * 1e-3) / 1e9, 1)})
    set_timing(e, False)
    e.close()
    print(json.dumps({"kernel": "k_copy", "bytes_read_per_launch": mib << 20, "bytes_written_per_launch": mib << 20,
                      "launches": n, "avg_launch_us": round(ms / n * 1e3, 2),
                      "gb_per_s": round(b / (ms * 1e-3) / 1e9, 1),
                      "read_kernel": {"bytes_read_per_launch": mib << 20, "launches": rn,
                                      "avg_launch_us": round(rms / rn * 1e3, 2),
                                      "gb_per_s": round(rb / (rms * 1e-3) / 1e9, 1)},
                      "gather_kernel": gather}))


def grid_case(args):
    """4096 x 4096 lattice, 4-neighbour, 3% of edges deleted (seed 3), edges of
    the largest component kept with the lattice ids (N = 2^24), -c 0 (the
    default adaptive stop c=1 x=1 holds after the first rounds here).  Two
    distributions: uniform (-z 0, exact batches of 256) and delta sampling
    (-z 4, batch-synchronous batches of 256).  With N = 2^24 the reference's
    retry cap (int)(N*N) is 0, so a duplicate draw consumes an attempt; the
    concentrated delta distribution makes most -z 4 attempts duplicates, so
    its pivots/s count rounds that actually ran.  Batches of 256 roots run as
    workgroup sweeps (one workgroup, i.e. one CU, per root); the bit-packed
    sweep (batches of 64, level loop from HIP graphs) is timed on the same
    -z 0 roots for comparison.  Plain rounds launch per level."""
    from axbc_amd import Engine, Graph, _ffi
    from axbc_amd.pivots import PivotLoop
    L = _ffi.hip()
    t = time.perf_counter()
    g = Graph.lattice(4096, 0.03, seed=3, heuristic=0)
    log(f"lattice: N={g.N} ned={g.ned} built in {time.perf_counter() - t:.1f}s")
    e = Engine(g)
    if args.no_graphs:
        if L.axbc_set_batch_graphs(e.ctx, 0) != 0:
            raise RuntimeError(L.axbc_last_error())
    out = {"case": "grid-4096-del3-c0", "n": g.N, "ned": g.ned}

    def timed(loop, steps):
        loop.step()  # warm-up (allocation, list growth, graph capture, sweep choice)
        e.sync()
        r0, a0 = loop.res.nrounds, loop.res.attempts
        ms0, n0, b0 = kernel_stats(e, 16)
        t = time.perf_counter()
        for _ in range(steps):
            if not loop.step():
                break
        e.sync()
        dt = time.perf_counter() - t
        ms1, n1, b1 = kernel_stats(e, 16)
        r = {"rounds": loop.res.nrounds - r0, "attempts": loop.res.attempts - a0, "seconds": round(dt, 2),
             "pivots_per_s": round((loop.res.nrounds - r0) / dt, 3) if dt > 0 else 0.0}
        if n1 > n0:
            r["k_wgsweep"] = {"launches": n1 - n0, "avg_launch_ms": round((ms1 - ms0) / (n1 - n0), 2),
                              "algorithmic_gb_per_s": round((b1 - b0) / ((ms1 - ms0) * 1e-3) / 1e9, 1)}
        return r

    for z in (0, 4):
        # the workgroup sweep at width 256: one root per CU
        assert L.axbc_set_batch_mode(e.ctx, 2) == 0
        assert L.axbc_set_batch_width(e.ctx, 256) == 0
        e.reset_bc()
        # events and byte counters on the workgroup sweep only (with timing
        # on, a bit-packed level loop launches from the host)
        set_timing_kinds(e, [16])
        b = timed(PivotLoop(e, 2048, c=0, x=1, z=z, batch=256), args.steps)
        set_timing_kinds(e, [])
        if z == 0:
            assert L.axbc_set_batch_width(e.ctx, 64) == 0
            assert L.axbc_set_batch_mode(e.ctx, 1) == 0
            e.reset_bc()
            bp = timed(PivotLoop(e, 2048, c=0, x=1, z=z, batch=64), max(2, args.steps // 6))
        e.reset_bc()
        p = timed(PivotLoop(e, 2048, c=0, x=1, z=z, batch=0), args.plain_rounds)
        out[f"z{z}"] = {"batched": b, "plain": p,
                        "batched_over_plain": round(b["pivots_per_s"] / p["pivots_per_s"], 2) if p["pivots_per_s"] else None}
        if z == 0:
            out["z0"]["bit_packed_width64"] = bp
        log(f"z{z}: {out[f'z{z}']}")
    st = e.pivot(int(np.flatnonzero(np.diff(g.col) > 0)[0]), 1.0 / g.N, stats=True)
    out["levels_of_a_round"] = st["nlevels"]
    e.close()
    print(json.dumps(out))


def serve_case(args):
    from axbc_amd.serve import poisson_jobs, run_workload
    jobs = poisson_jobs(args.serve, args.rate, seed=0)
    t = time.perf_counter()
    r = run_workload(jobs, slots=args.slots, hbm_cap=200e9)
    log(f"serving: {r} in {time.perf_counter() - t:.1f}s")
    case = f"serving-poisson{args.rate:g}-s12to20-z6"
    print(json.dumps(dict({"case": case, "requests": args.serve, "slots": args.slots, "rate": args.rate},
                          **{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()})))


def learned_case(args):
    """Accuracy on SCALE 18 (two generator seeds x three loop seeds): exact
    Brandes (every vertex a root, bit-packed batches) against the learned
    loop (with the engine's inputs and with the specification's: LCC and
    the last dependency row), and the uniform loop at 4096 pivots; then the
    learned loop's time on SCALE 24."""
    from axbc_amd import Engine, Graph
    from axbc_amd.learned import accuracy, exact_bc, run_learned
    from axbc_amd.pivots import PivotLoop
    out = {"case": "learned-vs-uniform-4096", "runs": []}
    for gseed in ((4, 5), (6, 7)):
        g = Graph.kronecker_device(18, 16, seed=gseed, heuristic=1)
        with Engine(g) as e:
            t = time.perf_counter()
            ex = exact_bc(e, batch=128)
            t_exact = time.perf_counter() - t
            for lseed in (4343543, 1, 2):
                info = run_learned(e, budget=4096, warm=512, batch=64, seed=lseed)
                a_l = accuracy(e.bc(), ex)
                info_s = run_learned(e, budget=4096, warm=512, batch=64, seed=lseed, features="spec")
                a_s = accuracy(e.bc(), ex)
                e.reset_bc()
                loop = PivotLoop(e, g.N, c=0, x=0, z=0, seed=lseed, batch=64)
                while loop.res.nrounds < 4096 and loop.step(min(64, 4096 - loop.res.nrounds)):
                    pass
                a_u = accuracy(e.bc(), ex)
                run = {"graph_seed": gseed, "loop_seed": lseed, "exact_seconds": round(t_exact, 2),
                       "learned": a_l, "learned_spec_inputs": a_s, "uniform": a_u, "fallback": info["fallback"],
                       "fallback_spec_inputs": info_s["fallback"],
                       "val_loss": [info["val_baseline"], info["val_loss"]],
                       "val_loss_spec_inputs": [info_s["val_baseline"], info_s["val_loss"]]}
                log(run)
                out["runs"].append(run)
    rl = [r["learned"]["top_recall"] for r in out["runs"]]
    ru = [r["uniform"]["top_recall"] for r in out["runs"]]
    sl = [r["learned"]["spearman"] for r in out["runs"]]
    su = [r["uniform"]["spearman"] for r in out["runs"]]
    rs_ = [r["learned_spec_inputs"]["top_recall"] for r in out["runs"]]
    ss_ = [r["learned_spec_inputs"]["spearman"] for r in out["runs"]]
    out.update({"top1pct_recall_learned": round(float(np.mean(rl)), 4), "top1pct_recall_uniform": round(float(np.mean(ru)), 4),
                "top1pct_recall_learned_spec_inputs": round(float(np.mean(rs_)), 4),
                "spearman_learned": round(float(np.mean(sl)), 5), "spearman_uniform": round(float(np.mean(su)), 5),
                "spearman_learned_spec_inputs": round(float(np.mean(ss_)), 5)})
    t = time.perf_counter()
    g = Graph.kronecker_device(24, 16, seed=(4, 5), heuristic=1)
    log(f"SCALE 24 built in {time.perf_counter() - t:.1f}s (N={g.N} ned={g.ned})")
    with Engine(g) as e:
        info = run_learned(e, budget=4096, warm=512, batch=64)
    out.update({"s24_learned_seconds": round(info["seconds"], 2), "s24_pivots_per_s": round(4096 / info["seconds"], 1),
                "s24_train_seconds": round(info["t_train"], 2), "s24_fallback": info["fallback"]})
    print(json.dumps(out))


def cpu_baseline(g, roots, probs, seconds):
    """The oracle on the same graph and root sequence, bounded to about
    `seconds` of wall time: half on one thread (the rate that sizes the
    rest), half with its rounds spread over the host's OpenMP threads
    (independent rounds of the static distribution), which is the value."""
    from oracle import oracle
    threads = max(1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    bc = np.zeros(g.N)
    done, er = 0, 0
    t0 = time.perf_counter()
    for v, p in zip(roots, probs):
        _, _, _, bc, st = oracle.pivot(g, int(v), float(p), bc)
        done += 1
        er += st["e_r"]
        if time.perf_counter() - t0 > seconds / 2:
            break
    dt1 = time.perf_counter() - t0
    rate1 = done / dt1
    m = int(min(len(roots) - done, max(threads, rate1 * threads * seconds / 2)))
    t0 = time.perf_counter()
    _, erm = oracle.pivots_parallel(g, roots[done:done + m], probs[done:done + m], threads)
    dtm = time.perf_counter() - t0
    return {"value": round(m / dtm, 4), "unit": "pivots/s", "cores": threads, "kind": "port",
            "sample": f"oracle/brandes_oracle.c on the same graph and roots: {m} pivots over {threads} OpenMP "
                      f"threads in {dtm:.1f} s (GTEPS {2 * erm / dtm / 1e9:.4f}); one thread: {done} pivots in "
                      f"{dt1:.1f} s ({rate1:.4f} pivots/s, GTEPS {2 * er / dt1 / 1e9:.4f}); host: {cpu_model()}, "
                      f"nproc {os.cpu_count()}"}


def cpu_model() -> str:
    """The host CPU's model name from /proc/cpuinfo (empty if unreadable)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24, help="timed batches")
    ap.add_argument("--warmup", type=int, default=2, help="untimed batches")
    ap.add_argument("--batch", type=int, default=0, choices=[0, 32, 64, 128, 256],
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
    ap.add_argument("--grid", action="store_true",
                    help="high-diameter case instead: 4096^2 lattice, 3%% edges deleted (seed 3), -z 0 and -z 4 "
                         "(batch-synchronous) in batches of 256 vs plain rounds")
    ap.add_argument("--serve", type=int, default=0, metavar="REQUESTS",
                    help="serving case instead: Poisson 20 req/s, SCALE 12-20, ef 8/16/32, budgets 64-2048, -z 6, "
                         "200 GB cap; prints p50/p99 latency and aggregate pivots/s")
    ap.add_argument("--rate", type=float, default=20.0, help="arrivals per second of --serve (the workload's is 20)")
    ap.add_argument("--slots", type=int, default=8,
                    help="device workers of --serve (two streams per hardware queue at HIP's 4 queues per process)")
    ap.add_argument("--learned", action="store_true",
                    help="learned-proposal case instead: accuracy vs exact Brandes on SCALE 18 (learned vs uniform at "
                         "4096 pivots) and the learned loop's throughput on SCALE 24")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the calibration copy kernel over MIB MiB (8 launches) and print its rate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs (pivot sharding with one "
                                   "all-reduce of bc) are not implemented", "n_gpus": args.gpus}))
        sys.exit(2)

    from axbc_amd import Engine, Graph
    from axbc_amd.pivots import PivotLoop

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.grid:
        grid_case(args)
        return
    if args.serve:
        serve_case(args)
        return
    if args.learned:
        learned_case(args)
        return

    t0 = time.perf_counter()
    g = Graph.kronecker_device(args.scale, 16, seed=(2, 3), heuristic=1)
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
        # roots from the workload's own distribution (high-degree neighbourhood),
        # drawn with another seed so that the tuning does not run the timed roots
        sample = PivotLoop(e, 4096, c=0, x=1024, z=1, seed=4343543 + 1).sample_roots(256 * 4)
        e.reset_bc()
        t = time.perf_counter()
        tuned = tune(e, sample, reps=3)  # 3 timed batches per setting: one batch alone is too noisy to rank them
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

    # the split threshold and pull ratio in force (the engine's defaults are 256 and 32)
    cur_chunk = tuned["chunk"] if tuned else (args.chunk or 256)
    cur_pull = tuned["pull"] if tuned else (args.pull if args.pull >= 0 else 32)

    # ---- batched pivot loop ---------------------------------------------
    # the warmup batches run with events on every kernel (the per-kernel
    # breakdown, which picks the dominant kernel); the timed batches carry
    # the events of the dominant kernel only, so the instrumentation of the
    # others (about 2 % of a batch) stays out of the measured rate
    loop = PivotLoop(e, 4096, c=1, x=1024, z=1, batch=batch)
    set_timing(e, True)
    for _ in range(args.warmup):
        loop.step()
    e.sync()
    wstats = {k: kernel_stats(e, k) for k in BATCH_KINDS}
    warm_ms = sum(v[0] for v in wstats.values())
    if warm_ms > 0:
        dom = max(wstats, key=lambda k: wstats[k][0])
        set_timing_kinds(e, [dom])
        log("per-kernel breakdown over the warmup batches:")
        log_kernels(wstats)
    else:
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
    if warm_ms <= 0:
        wstats, warm_ms = kstats, sum(v[0] for v in kstats.values())
        dom = max(kstats, key=lambda k: kstats[k][0])
        log_kernels(kstats)
    ms, n, b = kstats[dom]
    achieved = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(KIND_NAMES[dom], batch, cur_chunk, cur_pull),
                "kernel": KIND_NAMES[dom], "avg_launch_us": round(ms / max(n, 1) * 1e3, 2), "launches": n,
                "algorithmic_bytes_per_launch": int(b / max(n, 1)),
                "kernel_ms_share": round(wstats[dom][0] / max(warm_ms, 1e-9), 3)}
    if batch > 128:
        # batches of 256 run two 128-lane sweeps side by side, so each launch
        # shares the GPU with the other sweep's kernels and takes longer than
        # alone: the kernel's launches together over the time any of them ran
        # (overlapped), and the same kernel on batches of 128 of the
        # workload's roots after the timed region (alone)
        from axbc_amd import _ffi
        busy = C.c_double()
        if _ffi.hip().axbc_kernel_busy(e.ctx, dom, C.byref(busy)) != 0:
            raise RuntimeError(_ffi.hip().axbc_last_error())
        ach = b / (busy.value * 1e-3) / 1e9 if busy.value > 0 else 0.0
        roofline["overlapped"] = {"busy_ms": round(busy.value, 2), "summed_ms": round(ms, 2),
                                  "achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}
        roofline["alone"] = roofline_alone(e, dom, args.no_graphs)

    # ---- plain one-root-per-round loop on the same engine --------------------
    # the warm-up rounds carry events on every kernel (the breakdown); the
    # timed rounds carry none, as the batched loop's carry only its dominant
    # kernel's
    e.reset_bc()
    ploop = PivotLoop(e, 4096, c=0, x=1024, z=1, batch=0)
    set_timing(e, True)
    for _ in range(4):
        ploop.step()
    e.sync()
    log_kernels({k: kernel_stats(e, k) for k in PLAIN_KINDS})
    set_timing(e, False)
    r0p, er0p = ploop.res.nrounds, ploop.res.e_r
    t = time.perf_counter()
    for _ in range(args.plain_rounds):
        if not ploop.step():
            break
    e.sync()
    dtp = time.perf_counter() - t
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
