"""Pins the CPU reference of the row-size network by central finite
differences (the mathematics fixes the gradient)."""
import numpy as np

from oracle import mlp_ref


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n = 40
    w = rng.normal(0, 0.3, mlp_ref.H * mlp_ref.IN + mlp_ref.H + mlp_ref.H * mlp_ref.H + 2 * mlp_ref.H + 1)
    X = rng.uniform(0, 2, (n, mlp_ref.IN))
    ub = rng.uniform(10, 3000, n)
    y = ub * rng.uniform(0.05, 1.0, n)
    loss, g = mlp_ref.loss_and_grad(w, X, ub, y)
    idx = rng.choice(w.size, 60, replace=False)
    eps = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (mlp_ref.loss_and_grad(wp, X, ub, y)[0] - mlp_ref.loss_and_grad(wm, X, ub, y)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_reference_prediction_bounds():
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.3, 4673)
    X = rng.uniform(0, 2, (10, 6))
    ub = rng.uniform(1, 100, 10)
    p = mlp_ref.predict(w, X, ub)
    assert np.all(p > 0) and np.all(p < ub)
