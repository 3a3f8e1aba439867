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


def test_soft_target_gradient_matches_finite_differences():
    # Several labels per row, a label out of range (no label) and a row with none.
    rng = np.random.default_rng(2)
    dims = (6, 7, 5, 4)
    D, H1, H2, C = dims
    n = H1 * D + H1 + H2 * H1 + H2 + C * H2 + C
    params = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((9, D))
    labels = rng.integers(0, C, size=(9, 3))
    labels[2, 1] = 99
    labels[5, :] = 77
    loss, g = mlp_oracle.loss_grad(params, x, labels, dims, round16=False)
    idx = rng.choice(n, 40, replace=False)
    fd = mlp_oracle.finite_difference(params, x, labels, dims, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-6, atol=1e-9)
    # The row without a label contributes nothing: dropping it scales the loss by 9/8.
    keep = np.arange(9) != 5
    loss8, _ = mlp_oracle.loss_grad(params, x[keep], labels[keep], dims, round16=False)
    assert abs(loss * 9 - loss8 * 8) <= 1e-12 * abs(loss)


def test_soft_targets_reduce_to_one_hot():
    rng = np.random.default_rng(3)
    dims = (5, 6, 6, 7)
    D, H1, H2, C = dims
    n = H1 * D + H1 + H2 * H1 + H2 + C * H2 + C
    params = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((6, D))
    lab = rng.integers(0, C, size=6)
    a = mlp_oracle.loss_grad(params, x, lab, dims, round16=False)
    b = mlp_oracle.loss_grad(params, x, np.stack([lab, lab], 1), dims, round16=False)
    c = mlp_oracle.loss_grad(params, x, np.stack([lab, np.full(6, 1000)], 1), dims, round16=False)
    for o in (b, c):
        assert abs(o[0] - a[0]) <= 1e-12
        np.testing.assert_allclose(o[1], a[1], rtol=1e-12, atol=1e-15)
    # Two distinct labels: the loss is the mean of the two one-hot losses.
    lab2 = (lab + 1) % C
    l1, _ = mlp_oracle.loss_grad(params, x, lab2, dims, round16=False)
    l2, _ = mlp_oracle.loss_grad(params, x, np.stack([lab, lab2], 1), dims, round16=False)
    assert abs(l2 - 0.5 * (a[0] + l1)) <= 1e-12


def test_propose_lists_hand_example():
    # Softmax of log (4, 2, 2, 1, 1) = (0.4, 0.2, 0.2, 0.1, 0.1): equal logits go to the
    # lower index; 0.4 + 0.2 + 0.2 = 0.8 reaches 0.75 with 3 lists, and mass 0.95 needs all
    # five, which P = 4 does not hold.
    z = np.log(np.array([[4.0, 2.0, 2.0, 1.0, 1.0]]))
    lists, counts, _ = mlp_oracle.propose_lists(z, 4, 0.75)
    np.testing.assert_array_equal(lists[0], [0, 1, 2, 3])
    assert counts[0] == 3
    _, counts, _ = mlp_oracle.propose_lists(z, 4, 0.95)
    assert counts[0] == 0
    _, counts, _ = mlp_oracle.propose_lists(z, 1, 0.4)
    assert counts[0] == 1
