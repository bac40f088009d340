"""Learned pivot proposal: pivots drawn from an MLP's predicted betweenness.

A small MLP over per-vertex features (log degree, target adjacency, running
bc, mean neighbour bc) proposes pivots in place of the bc/delta heuristics.
It is trained on the method's own output: the first `warm` pivots are
uniform.  The bc estimate of the first quarter of them (the pilot) ranks the
vertices; the vertices around the top-1% boundary (ranks k/2 .. 2k), each
weighted by 1/bc, are the targets.  Target adjacency is the sum of those
weights over a vertex's neighbours (normalised by its mean): a root next to
boundary vertices sends many of its shortest paths through them (the LCC
takes its place when there are no targets).  For every later warm-up root, its
utility -- its share of the targets' estimates, (reach+1) *
sum_v delta_root(v) / bc(v), read from the batch's per-root delta rows on
the device -- is the training label, as log(1 + utility / mean).  Sampling
roots in proportion to that utility lowers the variance exactly where the
top-1% ranking is decided.  Training runs on the device (hand-written
forward/backward/Adam kernels, BF16 weights, FP32 accumulation) on three
quarters of those roots; the rest decide whether it generalises.  The
remaining pivots are drawn from w(v) = expm1(score(v)) + eps * mean -- a
static distribution, so they run in exact bit-packed batches; when the
held-out loss does not beat the mean predictor (or is not finite) the loop
falls back to uniform sampling.  Every round contributes delta/p with its
actual sampling probability p, so the estimator stays unbiased.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _ffi
from .engine import Engine, _check
from .pivots import PivotLoop, Rand48

N_PARAMS = 4 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    p[0:256] = rng.normal(0, np.sqrt(2.0 / 4), 256)
    o2 = 256 + 64
    p[o2:o2 + 4096] = rng.normal(0, np.sqrt(2.0 / 64), 4096)
    o3 = o2 + 4096 + 64
    p[o3:o3 + 64] = rng.normal(0, np.sqrt(1.0 / 64), 64)
    return p


def _hip():
    L = _ffi.hip()
    for name, args in {
        "axbc_mlp_init": [C.c_void_p, C.c_void_p],
        "axbc_mlp_features": [C.c_void_p, C.c_double],
        "axbc_mlp_features_spec": [C.c_void_p, C.c_double, C.c_int, C.c_double],
        "axbc_mlp_targets": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32],
        "axbc_mlp_set_data": [C.c_void_p, C.c_void_p, C.c_void_p],
        "axbc_mlp_get_data": [C.c_void_p, C.c_void_p, C.c_void_p],
        "axbc_mlp_gradient": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p],
        "axbc_mlp_adam": [C.c_void_p, C.c_float],
        "axbc_mlp_params": [C.c_void_p, C.c_void_p],
        "axbc_mlp_forward": [C.c_void_p],
        "axbc_mlp_proposal": [C.c_void_p, C.c_double, C.POINTER(C.c_uint64)],
    }.items():
        f = getattr(L, name)
        f.restype, f.argtypes = C.c_int, args
    return L


class Mlp:
    """Device MLP bound to an engine context."""

    def __init__(self, e: Engine, params: np.ndarray | None = None, seed: int = 0):
        self.e, self.L = e, _hip()
        p = np.ascontiguousarray(params if params is not None else init_params(seed), np.float32)
        _check(self.L.axbc_mlp_init(e.ctx, p.ctypes.data))

    def set_data(self, x=None, y=None):
        xa = None if x is None else np.ascontiguousarray(x, np.float32)
        ya = None if y is None else np.ascontiguousarray(y, np.float32)
        _check(self.L.axbc_mlp_set_data(self.e.ctx, None if xa is None else xa.ctypes.data,
                                        None if ya is None else ya.ctypes.data))

    def features(self, bc_scale: float):
        _check(self.L.axbc_mlp_features(self.e.ctx, bc_scale))

    def features_spec(self, bc_scale: float, k: int, delta_scale: float):
        """log(1+deg), LCC, log(1+bc*bc_scale), log(1+delta_k*delta_scale)."""
        _check(self.L.axbc_mlp_features_spec(self.e.ctx, bc_scale, int(k), delta_scale))

    def targets(self, vertex: np.ndarray, weight: np.ndarray):
        v = np.ascontiguousarray(vertex, np.uint32)
        w = np.ascontiguousarray(weight, np.float64)
        _check(self.L.axbc_mlp_targets(self.e.ctx, v.ctypes.data, w.ctypes.data, len(v)))

    def get(self):
        x = np.empty((self.e.N, 4), np.float32)
        s = np.empty(self.e.N, np.float32)
        _check(self.L.axbc_mlp_get_data(self.e.ctx, x.ctypes.data, s.ctypes.data))
        return x, s

    def gradient(self, idx) -> tuple[np.ndarray, float]:
        i = np.ascontiguousarray(idx, np.uint32)
        g = np.empty(N_PARAMS, np.float32)
        loss = C.c_float(0)
        _check(self.L.axbc_mlp_gradient(self.e.ctx, i.ctypes.data, len(i), g.ctypes.data, C.byref(loss)))
        return g, float(loss.value)

    def step(self, idx, lr: float) -> float:
        i = np.ascontiguousarray(idx, np.uint32)
        loss = C.c_float(0)
        _check(self.L.axbc_mlp_gradient(self.e.ctx, i.ctypes.data, len(i), None, C.byref(loss)))
        _check(self.L.axbc_mlp_adam(self.e.ctx, lr))
        return float(loss.value)

    def params(self) -> np.ndarray:
        p = np.empty(N_PARAMS, np.float32)
        _check(self.L.axbc_mlp_params(self.e.ctx, p.ctypes.data))
        return p

    def forward(self):
        _check(self.L.axbc_mlp_forward(self.e.ctx))

    def proposal(self, eps: float) -> int:
        t = C.c_uint64(0)
        _check(self.L.axbc_mlp_proposal(self.e.ctx, eps, C.byref(t)))
        return int(t.value)


class StaticLoop(PivotLoop):
    """Pivot loop whose draws come from the last fixed-point distribution
    built on the device (the proposal), batched."""

    def __init__(self, e, nbfs, total, rng: Rand48, batch=64):
        super().__init__(e, nbfs, c=0, x=0, z=0, batch=batch)
        self.rng = rng
        self._total = total

    def _draw(self):
        r = self._fixed_draw(self._total)
        v, p = C.c_uint32(0), C.c_double(0)
        _check(_ffi.hip().axbc_weights_pick(self.e.ctx, r, C.byref(v), C.byref(p)))
        return int(v.value), float(p.value)


def run_learned(e: Engine, budget: int = 4096, warm: int = 512, batch: int = 64, eps: float = 1.0,
                steps: int = 300, lr: float = 3e-3, seed: int = 4343543, top: float = 0.01,
                target_feature: bool = True, features: str = "targets") -> dict:
    """The learned loop; returns timing, training log and whether it fell
    back to uniform.  bc on the engine holds the estimate afterwards.
    features: "targets" (log degree, target adjacency, running bc, mean
    neighbour bc) or "spec" (log degree, LCC, running bc, the last
    warm-up root's dependency row)."""
    if features not in ("targets", "spec"):
        raise ValueError("features: 'targets' or 'spec'")
    t0 = time.perf_counter()
    g = e.g
    N = g.N
    L = _ffi.hip()
    e.reset_bc()
    pilot = max(1, warm // 4)
    loop = PivotLoop(e, N, c=0, x=0, z=0, seed=seed, batch=batch)
    n0 = 0  # first round of the last batch
    while loop.res.nrounds < pilot:
        n0 = loop.res.nrounds
        if not loop.step(min(batch, pilot - n0)):
            break
    bc_a = e.bc()
    # targets: the vertices around the top-k boundary of the pilot estimate
    k = max(1, int(round(top * N)))
    order = np.argsort(-bc_a, kind="stable")
    tv = order[k // 2:2 * k]
    tv = np.ascontiguousarray(tv[bc_a[tv] > 0], np.uint32)
    tw = np.ascontiguousarray(1.0 / bc_a[tv], np.float64)
    r1 = g.reach.astype(np.float64) + 1.0 if g.reach is not None else np.ones(N)
    util = {}
    out = np.zeros(256)  # one entry per lane of the widest batch
    while loop.res.nrounds < warm:
        n0 = loop.res.nrounds
        if not loop.step(min(batch, warm - n0)):
            break
        if len(tv):
            _check(L.axbc_batch_utility(e.ctx, tv.ctypes.data, tw.ctypes.data, len(tv), out.ctypes.data))
        for j, r in enumerate(loop.res.pivots[n0:loop.res.nrounds]):
            util[int(r)] = float(out[j]) * float(r1[r]) if len(tv) else 0.0
    bc_w = e.bc()
    t_warm = time.perf_counter() - t0
    use_targets = features == "targets" and target_feature and len(tv) > 0
    if not use_targets and features == "targets":
        _check(L.axbc_lcc_build(e.ctx, None, None, None, C.byref(C.c_uint64(0))))
    mlp = Mlp(e, seed=seed)
    mw = bc_w[bc_w > 0].mean() if (bc_w > 0).any() else 1.0
    if features == "spec":
        # the last warm-up root's dependency row, scaled by its mean
        k_last = max(0, loop.res.nrounds - 1 - n0)
        d_last = e.batch_delta(k_last)
        dm = d_last[d_last > 0].mean() if (d_last > 0).any() else 1.0
        mlp.features_spec(1.0 / mw, k_last, 1.0 / dm)
    else:
        if use_targets:
            mlp.targets(tv, tw)
        mlp.features(1.0 / mw)
    roots = np.array(sorted(util), np.uint32)
    u = np.array([util[int(r)] for r in roots])
    losses, ok, val, base = [], False, float("nan"), float("nan")
    if len(roots) >= 8 and (u > 0).any():
        mu = u[u > 0].mean()
        y = np.zeros(N, np.float32)
        y[roots] = np.log1p(u / mu)
        mlp.set_data(None, y)
        rs = np.random.default_rng(seed)
        perm = rs.permutation(len(roots))
        ntr = max(1, (3 * len(roots)) // 4)
        tr, va = roots[perm[:ntr]], roots[perm[ntr:]]
        for _ in range(steps):
            losses.append(mlp.step(tr, lr))
        mlp.forward()
        _, sc = mlp.get()
        val = float(np.mean((sc[va] - y[va]) ** 2))
        base = float(np.mean((y[va] - y[tr].mean()) ** 2))
        ok = bool(np.isfinite(val) and val < base)
    t_train = time.perf_counter() - t0 - t_warm
    rest = max(0, budget - loop.res.nrounds)
    if ok:
        total = mlp.proposal(eps)
        sl = StaticLoop(e, N, total, loop.rng, batch=batch)
        sl.vs = loop.vs
        sl.res.nrounds = 0
        while sl.res.nrounds < rest and sl.step(min(batch, rest - sl.res.nrounds)):
            pass
        nrounds = loop.res.nrounds + sl.res.nrounds
    else:
        while loop.res.nrounds < budget and loop.step(min(batch, budget - loop.res.nrounds)):
            pass
        nrounds = loop.res.nrounds
    e.sync()
    return {"nrounds": nrounds, "fallback": not ok, "samples": len(roots), "targets": int(len(tv)), "features": features,
            "loss_first": losses[0] if losses else None, "loss_final": float(np.mean(losses[-10:])) if losses else None,
            "val_loss": val, "val_baseline": base, "t_warm": t_warm, "t_train": t_train,
            "seconds": time.perf_counter() - t0}


def exact_bc(e: Engine, batch: int = 64) -> np.ndarray:
    """Exact Brandes: every vertex of positive degree as a root with p = 1,
    in bit-packed batches."""
    e.reset_bc()
    roots = np.flatnonzero(np.diff(e.g.col) > 0).astype(np.uint32)
    _check(_ffi.hip().axbc_set_batch_width(e.ctx, 128 if batch == 128 else 64))
    ones = np.ones(batch)
    for i in range(0, len(roots), batch):
        r = roots[i:i + batch]
        e.batch(r, ones[:len(r)])
    e.sync()
    return e.bc()


def accuracy(est: np.ndarray, exact: np.ndarray, top: float = 0.01) -> dict:
    """Top-`top` recall and Spearman rank correlation of an estimate."""
    from scipy.stats import spearmanr
    n = len(exact)
    k = max(1, int(round(top * n)))
    t_true = set(np.argsort(-exact, kind="stable")[:k].tolist())
    t_est = set(np.argsort(-est, kind="stable")[:k].tolist())
    rho = spearmanr(est, exact).statistic
    return {"top_recall": len(t_true & t_est) / k, "spearman": float(rho), "k": k}
