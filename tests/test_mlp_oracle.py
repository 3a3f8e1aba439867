"""Pins the FP64 proposal-network oracle: its analytic backward against central finite
differences (no FP16 rounding), and its forward against a direct evaluation."""
import numpy as np

import mlp_oracle


def test_analytic_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = (6, 7, 5, 4)
    D, H1, H2, C = dims
    n = H1 * D + H1 + H2 * H1 + H2 + C * H2 + C
    params = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((9, D))
    labels = rng.integers(0, C, size=9)
    _, g = mlp_oracle.loss_grad(params, x, labels, dims, round16=False)
    idx = rng.choice(n, 40, replace=False)
    fd = mlp_oracle.finite_difference(params, x, labels, dims, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-6, atol=1e-9)


def test_forward_matches_direct_evaluation():
    rng = np.random.default_rng(1)
    dims = (3, 4, 4, 3)
    D, H1, H2, C = dims
    n = H1 * D + H1 + H2 * H1 + H2 + C * H2 + C
    params = rng.standard_normal(n)
    x = rng.standard_normal((2, D))
    W1, b1, W2, b2, W3, b3 = mlp_oracle.split(params, *dims)
    import math
    g = lambda v: 0.5 * v * (1 + math.erf(v / math.sqrt(2)))
    for r in range(2):
        h1 = [g(sum(W1[i, j] * x[r, j] for j in range(D)) + b1[i]) for i in range(H1)]
        h2 = [g(sum(W2[i, j] * h1[j] for j in range(H1)) + b2[i]) for i in range(H2)]
        z3 = [sum(W3[i, j] * h2[j] for j in range(H2)) + b3[i] for i in range(C)]
        np.testing.assert_allclose(mlp_oracle.forward(params, x, dims, round16=False)["z3"][r], z3,
                                   rtol=1e-12)
