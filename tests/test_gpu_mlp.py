"""Learned proposal: the device MLP's loss and gradient against the float64
CPU oracle (oracle/mlp_oracle.py), training progress, and the learned loop
against uniform sampling at the same pivot budget, judged by exact Brandes."""
import numpy as np
import pytest

from axbc_amd import Engine, Graph
from axbc_amd.learned import Mlp, accuracy, exact_bc, init_params, run_learned
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def test_require(require_gpu):
    pass


def test_gradient_matches_oracle(require_gpu):
    g = Graph.kronecker(12, 8, seed=(1, 2))
    rng = np.random.default_rng(0)
    x = rng.normal(size=(g.N, 4)).astype(np.float32)
    y = rng.normal(size=g.N).astype(np.float32)
    p = init_params(3)
    with Engine(g) as e:
        m = Mlp(e, p)
        m.set_data(x, y)
        idx = rng.choice(g.N, size=1000, replace=False).astype(np.uint32)
        grad, loss = m.gradient(idx)
        m.forward()
        _, score = m.get()
    pb = mo.to_bf16(p).astype(np.float64)
    loss_ref, grad_ref = mo.loss_and_grad(pb, x[idx].astype(np.float64), y[idx].astype(np.float64))
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    scale = np.abs(grad_ref).max()
    assert np.abs(grad - grad_ref).max() <= 1e-4 * scale
    s_ref, _ = mo.forward(pb, x.astype(np.float64))
    assert np.abs(score - s_ref).max() <= 1e-4 * (np.abs(s_ref).max() + 1)


def test_training_reduces_loss(require_gpu):
    g = Graph.kronecker(12, 16, seed=(1, 2))
    rng = np.random.default_rng(1)
    x = rng.normal(size=(g.N, 4)).astype(np.float32)
    y = (np.tanh(x[:, 0]) + 0.5 * x[:, 1] ** 2).astype(np.float32)
    with Engine(g) as e:
        m = Mlp(e, init_params(0))
        m.set_data(x, y)
        losses = [m.step(rng.choice(g.N, 2048, replace=False), 3e-3) for _ in range(300)]
    assert np.mean(losses[-20:]) < 0.2 * np.mean(losses[:5])


def test_learned_vs_uniform(require_gpu):
    g = Graph.kronecker(13, 16, seed=(4, 5), heuristic=1)
    with Engine(g) as e:
        ex = exact_bc(e)
        info = run_learned(e, budget=512, warm=128, batch=64, steps=150, minibatch=4096)
        est_l = e.bc()
        e.reset_bc()
        from axbc_amd.pivots import PivotLoop
        loop = PivotLoop(e, g.N, c=0, x=0, z=0, batch=64)
        while loop.res.nrounds < 512 and loop.step(min(64, 512 - loop.res.nrounds)):
            pass
        est_u = e.bc()
    a_l, a_u = accuracy(est_l, ex), accuracy(est_u, ex)
    print("learned", info, a_l, "uniform", a_u, flush=True)
    assert info["nrounds"] == 512
    assert a_l["spearman"] > 0.5 and a_u["spearman"] > 0.5
    assert a_l["top_recall"] >= a_u["top_recall"] - 0.1
