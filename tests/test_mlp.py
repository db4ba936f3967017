"""The learned screen's network (include/xbach.h XB_MLP_*, csrc/k_screen.hip)
against its CPU restatement (oracle/mlp_oracle.py).

CPU: the oracle's gradient equals central finite differences of its loss
(float64, FP32-valued weights so the loss is smooth) to 1e-6 relative — this
pins the restatement by the mathematics, not by retyping it.
GPU: the HIP forward pass within 2e-6 of the oracle's probabilities; the HIP
gradient (FP32 accumulation over 777 samples, FP16 weights) within 1e-4 of the
largest oracle gradient component, and its loss within 1e-5 relative; training
(Adam, seeded) is deterministic and separates a toy problem.
"""
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import mlp_oracle as mo  # noqa: E402
import xbach  # noqa: E402


def random_model(rng, fp16_exact=False):
    m = np.zeros(mo.NPARAM, np.float32)
    m[mo.MU:mo.SD] = rng.normal(0, 1, 32)
    m[mo.SD:mo.W1] = rng.uniform(0.5, 2.0, 32)
    m[mo.W1:mo.B1] = rng.normal(0, 0.3, 64 * 32)
    m[mo.B1:mo.W2] = rng.normal(0, 0.1, 64)
    m[mo.W2:mo.B2] = rng.normal(0, 0.3, 64)
    m[mo.B2] = 0.1
    m[mo.TAU] = 0.5
    if fp16_exact:
        m[mo.W1:mo.B2] = m[mo.W1:mo.B2].astype(np.float16).astype(np.float32)
    return m


def data(rng, n):
    X = rng.normal(0, 1.5, (n, 32)).astype(np.float32)
    y = (rng.uniform(size=n) < 0.3).astype(np.float32)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return X, y, wt


def test_oracle_gradient_equals_finite_differences():
    rng = np.random.default_rng(3)
    m = random_model(rng).astype(np.float64)
    X, y, wt = data(rng, 57)
    g = mo.grad(m, X, y, wt, fp16=False)
    idx = np.concatenate([rng.choice(64 * 32, 40, replace=False),
                          64 * 32 + rng.choice(129, 20, replace=False)])
    h = 1e-6
    for q in idx:
        mp = m.copy()
        mn = m.copy()
        mp[mo.W1 + q] += h
        mn[mo.W1 + q] -= h
        fd = (mo.loss(mp, X, y, wt, fp16=False) - mo.loss(mn, X, y, wt, fp16=False)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, np.abs(g).max()), (q, fd, g[q])


def test_oracle_probabilities_and_loss_are_consistent():
    rng = np.random.default_rng(4)
    m = random_model(rng)
    X, y, wt = data(rng, 100)
    p, z, _, _ = mo.forward(m, X)
    assert np.all((p > 0) & (p < 1))
    ref = -np.sum(wt * (y * np.log(p) + (1 - y) * np.log1p(-p))) / len(y)
    assert abs(mo.loss(m, X, y, wt) - ref) <= 1e-9 * abs(ref)


@pytest.fixture(scope="module")
def ctx():
    c = xbach.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(ctx):
    rng = np.random.default_rng(5)
    m = random_model(rng)
    X, _, _ = data(rng, 300)
    p = ctx.mlp_forward(m, X)
    pr, _, _, _ = mo.forward(m, X)
    assert np.max(np.abs(p - pr)) <= 2e-6


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(ctx):
    rng = np.random.default_rng(6)
    m = random_model(rng)
    X, y, wt = data(rng, 777)  # not a multiple of the 64-sample chunk
    g, l = ctx.mlp_grad(m, X, y, wt)
    gr = mo.grad(m, X, y, wt)
    lr = mo.loss(m, X, y, wt)
    scale = np.abs(gr).max()
    assert np.max(np.abs(g - gr)) <= 1e-4 * scale, np.max(np.abs(g - gr)) / scale
    assert abs(l - lr) <= 1e-5 * abs(lr)


@pytest.mark.gpu
def test_gpu_training_is_deterministic_and_learns(ctx):
    rng = np.random.default_rng(7)
    n = 4000
    X = rng.normal(0, 1, (n, 32)).astype(np.float32)
    y = (X[:, 0] + 0.5 * X[:, 3] - 0.3 * X[:, 7] > 0.2).astype(np.float32)
    wt = np.ones(n, np.float32)
    m0 = random_model(rng)
    m0[mo.MU:mo.SD] = 0.0
    m0[mo.SD:mo.W1] = 1.0
    m1, loss1 = ctx.mlp_train(m0, X, y, wt, epochs=30, batch=256, lr=1e-2, seed=11)
    m2, loss2 = ctx.mlp_train(m0, X, y, wt, epochs=30, batch=256, lr=1e-2, seed=11)
    assert np.array_equal(m1, m2) and np.array_equal(loss1, loss2)
    assert loss1[-1] < 0.25 * loss1[0]
    acc = np.mean((ctx.mlp_forward(m1, X) > 0.5) == (y > 0.5))
    assert acc > 0.97, acc
