"""Pins of the CPU scorer restatement (oracle/mlp.py): BF16 rounding on
known values and the analytic gradient against central finite differences."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp  # noqa: E402


def test_bf16_round_to_nearest_even():
    e = 2.0 ** -8
    cases = {1.0: 1.0, 1 + e: 1.0, 1 + 3 * e: 1 + 4 * e, 1 + e + 2.0 ** -20: 1 + 2 * e,
             -(1 + e): -1.0, 3.0: 3.0, 0.0: 0.0, 1 + 2 * e: 1 + 2 * e}
    x = np.array(list(cases), dtype=np.float32)
    want = np.array(list(cases.values()), dtype=np.float32)
    assert np.array_equal(mlp.bf16(x), want)
    r = mlp.bf16(np.random.default_rng(0).normal(size=1000).astype(np.float32))
    assert np.all(r.view(np.uint32) & 0xFFFF == 0)


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    p = mlp.init_like(seed).astype(np.float64)
    x = rng.uniform(0, 1, (9, mlp.FEATURES))
    t = rng.uniform(0, 1, 9)
    g, loss = mlp.grad(p, x, t, bf16_mode=False)
    y = mlp.forward(p, x, bf16_mode=False)
    assert loss == pytest.approx(np.mean((y - t) ** 2))
    groups = [(mlp.P_W1, mlp.P_B1), (mlp.P_B1, mlp.P_W2), (mlp.P_W2, mlp.P_B2), (mlp.P_B2, mlp.P_W3),
              (mlp.P_W3, mlp.P_B3), (mlp.P_B3, mlp.P_N)]
    idx = []
    for a, b in groups:
        idx += list(rng.choice(np.arange(a, b), size=min(12, b - a), replace=False))
    h = 1e-6
    for i in idx:
        if mlp.P_W1 <= i < mlp.P_B1 and (i - mlp.P_W1) % 32 >= mlp.FEATURES:
            assert g[i] == 0.0  # padded input columns
            continue
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = np.mean((mlp.forward(pp, x, bf16_mode=False) - t) ** 2)
        lm = np.mean((mlp.forward(pm, x, bf16_mode=False) - t) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_bf16_mode_is_close_to_exact():
    rng = np.random.default_rng(3)
    p = mlp.init_like(3)
    x = rng.uniform(0, 1, (200, mlp.FEATURES)).astype(np.float32)
    t = rng.uniform(0, 1, 200).astype(np.float32)
    g16, l16 = mlp.grad(p, x, t, bf16_mode=True)
    g64, l64 = mlp.grad(p.astype(np.float64), x, t, bf16_mode=False)
    assert np.linalg.norm(g16 - g64) <= 0.03 * np.linalg.norm(g64)
    assert l16 == pytest.approx(l64, rel=0.02)


def test_adam_first_step_moves_by_lr():
    g = np.array([1e-3, -2.0, 0.0, 5e-6])
    p, m, v = mlp.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01)
    assert np.allclose(p[:2], [-0.01, 0.01], rtol=1e-4)
    assert p[2] == 0.0
