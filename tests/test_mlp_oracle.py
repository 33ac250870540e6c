"""Pins the float64 oracle of the first-substep network (oracle/mlp_oracle.py):
GELU closed-form values, the loss gradient against central finite differences,
and Adam's first step, whose bias correction gives lr * g / (|g| + eps)."""
import numpy as np

import mlp_oracle as M


def test_gelu_known_values():
    assert M.gelu(np.array(0.0)) == 0.0
    assert abs(M.gelu(np.array(10.0)) - 10.0) < 1e-12          # tanh saturates to 1
    assert abs(M.gelu(np.array(-10.0))) < 1e-12
    assert abs(M.gelu_d(np.array(0.0)) - 0.5) < 1e-15
    # tanh form at x = 1: 0.5 * (1 + tanh(sqrt(2/pi) * 1.044715))
    ref = 0.5 * (1.0 + np.tanh(np.sqrt(2.0 / np.pi) * 1.044715))
    assert abs(M.gelu(np.array(1.0)) - ref) < 1e-15
    x = np.linspace(-4, 4, 41)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_d(x))) < 1e-8


def test_loss_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, M.NPARAM)
    X = rng.normal(0, 1, (M.NIN, 7))
    t = rng.normal(0, 1, 7)
    L, g = M.loss_and_grad(p, X, t, n_total=10)
    # loss closed form
    y, _ = M.forward(p, X)
    assert abs(L - np.sum((y - t) ** 2) / 10) < 1e-14
    idx = np.concatenate([rng.choice(np.arange(0, 1536), 12, replace=False), [1536, 1599],
                          rng.choice(np.arange(1600, 5696), 12, replace=False), [5696, 5759],
                          [5760, 5800, 5823, 5824]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, X, t, 10)[0] - M.loss_and_grad(pm, X, t, 10)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * (1 + abs(g[i])), (i, fd, g[i])


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, m, v = M.adam(p, np.zeros(50), np.zeros(50), g, 1, lr=1e-2, eps=1e-8)
    np.testing.assert_allclose(p - p1, 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-15)
    np.testing.assert_allclose(v, 0.001 * g * g, rtol=1e-12)
