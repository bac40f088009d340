"""Learned pivot proposal (the north star's third pattern).

A small MLP over per-vertex features (log degree, LCC, running bc, last
delta) proposes pivots in place of the bc/delta heuristics.  It is trained
on the method's own output: the first `warm` pivots are uniform; the bc
estimate of the first half of them gives the features and the estimate of
the second half the target (cross-fitting, so the network learns to denoise
the running bc with degree and LCC rather than to copy it).  Training runs
on the device (hand-written forward/backward/Adam kernels, BF16 weights,
FP32 accumulation).  The remaining pivots are drawn from
w(v) = softplus(score(v)) + eps * mean softplus -- a static distribution, so
they run in exact bit-packed batches; when training fails (non-finite or
not improving loss) the loop falls back to uniform sampling.  Every round
contributes delta/p with its actual sampling probability p, so the
estimator stays unbiased.
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
        "axbc_mlp_features": [C.c_void_p, C.c_double, C.c_double],
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

    def features(self, bc_scale: float, delta_scale: float):
        _check(self.L.axbc_mlp_features(self.e.ctx, bc_scale, delta_scale))

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


def run_learned(e: Engine, budget: int = 4096, warm: int = 512, batch: int = 64, eps: float = 0.2,
                steps: int = 300, minibatch: int = 16384, lr: float = 3e-3, seed: int = 4343543) -> dict:
    """The learned loop; returns timing, training log and whether it fell
    back to uniform.  bc on the engine holds the estimate afterwards."""
    t0 = time.perf_counter()
    g = e.g
    N = g.N
    e.reset_bc()
    half = warm // 2
    loop = PivotLoop(e, N, c=0, x=0, z=0, seed=seed, batch=batch)
    while loop.res.nrounds < half and loop.step(min(batch, half - loop.res.nrounds)):
        pass
    bc_a = e.bc()
    while loop.res.nrounds < warm and loop.step(min(batch, warm - loop.res.nrounds)):
        pass
    bc_w = e.bc()
    bc_b = bc_w - bc_a
    t_warm = time.perf_counter() - t0
    # features from the first half, target from the second half
    _check(_ffi.hip().axbc_lcc_build(e.ctx, None, None, None, C.byref(C.c_uint64(0))))
    mlp = Mlp(e, seed=seed)
    ma = bc_a[bc_a > 0].mean() if (bc_a > 0).any() else 1.0
    mb = bc_b[bc_b > 0].mean() if (bc_b > 0).any() else 1.0
    e.set_bc(bc_a)
    mlp.features(1.0 / ma, 1.0 / max(float(e.delta().mean()), 1e-30))
    e.set_bc(bc_w)
    y = np.log1p(bc_b / mb).astype(np.float32)
    mlp.set_data(None, y)
    cand = np.flatnonzero(np.diff(g.col) > 0).astype(np.uint32)
    rs = np.random.default_rng(seed)
    losses = []
    for s in range(steps):
        idx = rs.choice(cand, size=min(minibatch, len(cand)), replace=False)
        losses.append(mlp.step(idx, lr))
    # baseline loss: predicting the target mean
    base = float(((y[cand] - y[cand].mean()) ** 2).mean())
    final = float(np.mean(losses[-10:]))
    ok = bool(np.isfinite(final) and final < base)
    t_train = time.perf_counter() - t0 - t_warm
    # inference on the features of the whole warm-up estimate
    mw = bc_w[bc_w > 0].mean() if (bc_w > 0).any() else 1.0
    mlp.features(1.0 / mw, 1.0 / max(float(e.delta().mean()), 1e-30))
    rest = max(0, budget - loop.res.nrounds)
    if ok:
        mlp.forward()
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
    return {"nrounds": nrounds, "fallback": not ok, "loss_first": losses[0] if losses else None,
            "loss_final": final, "loss_baseline": base, "t_warm": t_warm, "t_train": t_train,
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
