"""Pinning of the FP64 MLP oracle by central finite differences, and of its
Adam step by the closed form of the first step (|update| = lr per weight)."""
import numpy as np
import pytest

import mlp_ref


@pytest.mark.parametrize("dims", [(3, 5, 4, 2), (14, 8, 6, 5)])
def test_gradients_match_finite_differences(dims):
    rng = np.random.default_rng(sum(dims))
    _, n = mlp_ref.layout(dims)
    p = rng.normal(0, 0.5, n)
    X = rng.normal(0, 1, (7, dims[0]))
    T = rng.normal(0, 1, (7, dims[3]))
    M = (rng.random((7, dims[3])) < 0.8).astype(np.float64)
    g = mlp_ref.gradients(p, dims, X, T, M)
    h = 1e-6
    for k in rng.choice(n, size=min(n, 60), replace=False):
        e = np.zeros(n)
        e[k] = h
        fd = (mlp_ref.loss(p + e, dims, X, T, M) - mlp_ref.loss(p - e, dims, X, T, M)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), k


def test_first_adam_step_closed_form():
    rng = np.random.default_rng(0)
    g = rng.normal(0, 1, 50)
    p = rng.normal(0, 1, 50)
    p2, m1, m2 = mlp_ref.adam(p, g, np.zeros(50), np.zeros(50), 1, 0.01)
    # bias-corrected first step: m / sqrt(v) = sign(g) (up to eps)
    np.testing.assert_allclose(p - p2, 0.01 * np.sign(g), rtol=1e-6)


def test_loss_zero_at_target():
    dims = (4, 3, 3, 2)
    rng = np.random.default_rng(1)
    _, n = mlp_ref.layout(dims)
    p = rng.normal(0, 1, n)
    X = rng.normal(0, 1, (5, 4))
    _, acts = mlp_ref.forward(p, dims, X)
    assert mlp_ref.loss(p, dims, X, acts[-1]) == 0.0
    assert np.abs(mlp_ref.gradients(p, dims, X, acts[-1])).max() == 0.0
