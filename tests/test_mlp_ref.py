"""Pins the CPU oracle of the cost model: analytic gradient against central
finite differences, and the parameter layout against a hand-built network."""
import numpy as np

import mlp_ref as R


def test_gradient_matches_finite_differences():
    C = 5
    rng = np.random.default_rng(0)
    n_params = 16 * 64 + 64 + 64 * 64 + 64 + C * 64 + C
    p = rng.normal(scale=0.3, size=n_params)
    x = rng.random((40, 16))
    t = rng.normal(size=(40, C))
    m = (rng.random((40, C)) > 0.3).astype(float)
    w = rng.uniform(0.5, 2, 40)
    _, g = R.loss_grad(p, C, x, t, m, w)
    idx = list(rng.choice(n_params, 120, replace=False)) + [n_params - 1, 0, 16 * 64]
    fd = R.finite_difference(p, C, x, t, m, w, idx)
    for i in idx:
        assert abs(g[i] - fd[i]) <= 1e-6 + 1e-5 * abs(fd[i]), i


def test_layout_against_hand_built_network():
    C = 2
    W1 = np.zeros((64, 16)); W1[0, 3] = 2.0; W1[1, 4] = -1.0
    b1 = np.zeros(64); b1[1] = 5.0
    W2 = np.zeros((64, 64)); W2[7, 0] = 1.0; W2[8, 1] = 0.5
    b2 = np.zeros(64)
    W3 = np.zeros((C, 64)); W3[0, 7] = 1.0; W3[1, 8] = 3.0
    b3 = np.array([0.25, -1.0])
    p = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])
    x = np.zeros((1, 16)); x[0, 3] = 1.5; x[0, 4] = 2.0
    # h1[0] = 3, h1[1] = relu(-2 + 5) = 3; h2[7] = 3, h2[8] = 1.5; y = [3.25, 3.5]
    np.testing.assert_allclose(R.forward(p, C, x), [[3.25, 3.5]])
