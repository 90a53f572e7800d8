"""Pins the numpy cost-model reference (oracle/mlp.py) that checks the HIP
network kernels: finite differences for the gradient, closed forms for the
forward pass."""
import numpy as np

from oracle import mlp


def test_layout_and_parameter_count():
    # 12 -> 64 -> 64 -> 64 -> 1: 12*64+64 + 2*(64*64+64) + 64+1
    assert mlp.param_count(12, 64, 3) == 9217
    lay = mlp.layout(12, 64, 3)
    assert lay[0] == (0, 12, 64, 768)
    assert lay[-1][1:3] == (64, 1)


def test_forward_closed_form():
    # one hidden layer of 2 units, hand-set weights
    inputs, hidden, layers = 2, 2, 1
    p = np.zeros(mlp.param_count(inputs, hidden, layers))
    w0, a, b, b0 = mlp.layout(inputs, hidden, layers)[0]
    p[w0:w0 + 4] = [1.0, -1.0, 2.0, 0.5]  # W[0]=[1,-1], W[1]=[2,0.5]
    p[b0:b0 + 2] = [0.0, 1.0]
    w1, _, _, b1 = mlp.layout(inputs, hidden, layers)[1]
    p[w1:w1 + 2] = [3.0, -2.0]
    p[b1] = 0.25
    x = np.array([[1.0, 1.0], [-1.0, 0.0]])
    # row 0: z = [1+2, -1+0.5+1] = [3, 0.5] -> y = 9 - 1 + 0.25
    # row 1: z = [-1, 1+1] = [-1, 2] -> relu [0, 2] -> y = -4 + 0.25
    y, _ = mlp.forward(p, x, inputs, hidden, layers)
    np.testing.assert_allclose(y, [8.25, -3.75])


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    inputs, hidden, layers = 5, 7, 3
    p = rng.standard_normal(mlp.param_count(inputs, hidden, layers)) * 0.5
    x = rng.standard_normal((9, inputs))
    t = rng.standard_normal(9)
    loss, g = mlp.loss_grad(p, x, t, inputs, hidden, layers)
    h = 1e-6
    for k in rng.choice(len(p), 60, replace=False):
        e = np.zeros_like(p)
        e[k] = h
        lp, _ = mlp.loss_grad(p + e, x, t, inputs, hidden, layers)
        lm, _ = mlp.loss_grad(p - e, x, t, inputs, hidden, layers)
        assert abs((lp - lm) / (2 * h) - g[k]) < 1e-6 * max(1.0, abs(g[k])), k


def test_adam_reduces_a_quadratic():
    p = np.array([3.0, -2.0])
    m = np.zeros(2)
    v = np.zeros(2)
    for step in range(1, 501):
        p, m, v = mlp.adam_step(p, 2 * p, m, v, step, 0.05)
    assert np.all(np.abs(p) < 0.05)
