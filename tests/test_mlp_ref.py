"""Pins the float64 MLP reference (tests/mlp_ref.py) by central finite differences."""
import numpy as np

import mlp_ref


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    n_in, H, n = 5, 7, 23
    p = rng.normal(0, 0.5, mlp_ref.num_params(n_in, H))
    X = rng.normal(0, 1, (n, n_in))
    y = rng.normal(0, 1, n)
    loss, g = mlp_ref.loss_grad(p, X, y, n_in, H)
    eps = 1e-6
    num = np.zeros_like(p)
    for i in range(p.size):
        q = p.copy(); q[i] += eps
        lp, _ = mlp_ref.loss_grad(q, X, y, n_in, H)
        q[i] -= 2 * eps
        lm, _ = mlp_ref.loss_grad(q, X, y, n_in, H)
        num[i] = (lp - lm) / (2 * eps)
    np.testing.assert_allclose(g, num, rtol=1e-5, atol=1e-7)


def test_reference_forward_known_values():
    # one input, one hidden unit: pred = w3 * relu(w2 * relu(w1 x + b1) + b2) + b3
    p = np.array([2.0, 0.5, 3.0, -1.0, 4.0, 0.25])
    pred, _ = mlp_ref.forward(p, np.array([[1.5]]), 1, 1)
    assert pred[0] == 4.0 * max(3.0 * max(2.0 * 1.5 + 0.5, 0) - 1.0, 0) + 0.25
