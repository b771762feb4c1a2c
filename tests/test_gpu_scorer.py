"""Scorer kernels (BF16 MLP on the matrix cores) against the CPU restatement
oracle/mlp.py: forward, gradient, Adam step and training.

Tolerances: the CPU side rounds to BF16 at the same points as the device, so
the remaining differences come from FP32 summation order (and the rare BF16
rounding flip it causes): scores within 2e-3 absolute + 2e-3 relative,
gradient blocks within 1e-2 of their L2 norm.  Against the exact float64
gradient (no BF16) the device is within 3e-2 of the norm."""
import numpy as np
import pytest

import mlp

pytestmark = pytest.mark.gpu

GROUPS = {"W1": (mlp.P_W1, mlp.P_B1), "b1": (mlp.P_B1, mlp.P_W2), "W2": (mlp.P_W2, mlp.P_B2),
          "b2": (mlp.P_B2, mlp.P_W3), "w3": (mlp.P_W3, mlp.P_B3), "b3": (mlp.P_B3, mlp.P_N)}


@pytest.fixture(scope="module")
def scorer(ltl):
    sc = ltl.Scorer(seed=7)
    yield sc
    sc.close()


def _data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (n, mlp.FEATURES)).astype(np.float32)
    x[:, 12:19] = 0
    x[np.arange(n), 12 + rng.integers(0, 7, n)] = 1  # operator one-hot
    x[:, 23] = 1
    t = rng.uniform(0, 1, n).astype(np.float32)
    return x, t


def test_initial_params(ltl):
    a, b, c = ltl.Scorer(seed=7), ltl.Scorer(seed=7), ltl.Scorer(seed=8)
    p = a.params()
    assert p.shape == (mlp.P_N,) and np.all(np.isfinite(p))
    W1 = p[mlp.P_W1:mlp.P_B1].reshape(128, 32)
    assert np.all(W1[:, mlp.FEATURES:] == 0)
    assert np.array_equal(p, b.params())
    assert not np.array_equal(c.params(), p)
    for s in (a, b, c):
        s.close()


@pytest.mark.parametrize("n", [1, 16, 300, 4097])
def test_forward_matches_cpu(scorer, n):
    p = mlp.init_like(11)
    scorer.set_params(p)
    x, _ = _data(n, n)
    y = scorer.forward(x)
    yc = mlp.forward(p, x)
    assert np.all(np.abs(y - yc) <= 2e-3 + 2e-3 * np.abs(yc)), np.max(np.abs(y - yc))


@pytest.mark.parametrize("n,seed", [(64, 1), (300, 2), (1000, 3), (5000, 4)])
def test_grad_matches_cpu(scorer, n, seed):
    p = mlp.init_like(seed)
    scorer.set_params(p)
    x, t = _data(seed + 100, n)
    g, loss = scorer.grad(x, t)
    gc, lc = mlp.grad(p, x, t)
    g64, _ = mlp.grad(p.astype(np.float64), x, t, bf16_mode=False)
    assert loss == pytest.approx(lc, rel=1e-3)
    for name, (a, b) in GROUPS.items():
        ref = gc[a:b]
        err = np.linalg.norm(g[a:b] - ref)
        assert err <= 1e-2 * np.linalg.norm(ref) + 1e-7, (name, err, np.linalg.norm(ref))
    assert np.linalg.norm(g - g64) <= 3e-2 * np.linalg.norm(g64)
    W1g = g[mlp.P_W1:mlp.P_B1].reshape(128, 32)
    assert np.all(W1g[:, mlp.FEATURES:] == 0)


def test_adam_step_matches_cpu(ltl, scorer):
    p0 = mlp.init_like(5)
    scorer.set_params(p0)
    x, t = _data(5, 700)
    rows = np.concatenate([x, t[:, None]], axis=1)
    lr = 1e-3
    scorer.train(rows, epochs=1, batch=len(rows), lr=lr, seed=3)  # one step over every row
    p1 = scorer.params()
    gc, _ = mlp.grad(p0, x, t)
    want, _, _ = mlp.adam_step(p0.astype(np.float64), gc.astype(np.float64), 0.0, 0.0, 1, lr)
    big = np.abs(gc) > 1e-3 * np.abs(gc).max()
    assert np.all(np.abs(p1[big] - want[big]) <= 0.05 * lr)
    assert np.all(np.abs(p1 - p0) <= 1.001 * lr)


def test_training_reduces_loss(ltl):
    sc = ltl.Scorer(seed=3)
    x, _ = _data(9, 20000)
    t = (0.7 * x[:, 0] * x[:, 6] + 0.3 * (x[:, 2] < 0.2)).astype(np.float32)
    rows = np.concatenate([x, t[:, None]], axis=1)
    _, l0 = sc.grad(x, t)
    l1 = sc.train(rows, epochs=30, batch=512, lr=2e-3, seed=1)
    _, l2 = sc.grad(x, t)
    sc.close()
    assert l2 < 0.2 * l0, (l0, l1, l2)
