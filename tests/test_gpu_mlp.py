"""The selector's FP32 MLP kernels against the FP64 oracle (oracle/mlp_ref.py,
itself pinned by finite differences in tests/test_mlp.py): forward,
gradients of the masked MSE, the Adam update, and a training run."""
import numpy as np
import pytest

from spmvdare import hip
import mlp_ref

pytestmark = pytest.mark.gpu
DIMS = (14, 64, 64, 40)


@pytest.fixture(scope="module", autouse=True)
def device():
    hip.init(0)


def data(seed, rows=96):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (rows, DIMS[0])).astype(np.float32)
    T = rng.normal(0, 1, (rows, DIMS[3])).astype(np.float32)
    M = (rng.random((rows, DIMS[3])) < 0.9).astype(np.float32)
    return X, T, M


def test_forward_matches_oracle():
    net = hip.MLP(*DIMS, seed=3)
    X, _, _ = data(1)
    p = net.get_params().astype(np.float64)
    _, acts = mlp_ref.forward(p, DIMS, X.astype(np.float64))
    Y = net.predict(X)
    np.testing.assert_allclose(Y, acts[-1], rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("masked", [False, True])
def test_gradients_match_oracle(masked):
    net = hip.MLP(*DIMS, seed=5)
    X, T, M = data(2)
    M = M if masked else None
    g, loss = net.gradients(X, T, M)
    p = net.get_params().astype(np.float64)
    g_ref = mlp_ref.gradients(p, DIMS, X.astype(np.float64), T.astype(np.float64),
                              None if M is None else M.astype(np.float64))
    l_ref = mlp_ref.loss(p, DIMS, X.astype(np.float64), T.astype(np.float64),
                         None if M is None else M.astype(np.float64))
    assert loss == pytest.approx(l_ref, rel=1e-4)
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print("gradient relative error (FP32 device vs FP64 oracle): %.2e" % err)
    assert err < 1e-4


def test_adam_step_matches_oracle():
    net = hip.MLP(*DIMS, seed=7)
    rng = np.random.default_rng(4)
    p = net.get_params().astype(np.float64)
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    for step in range(1, 4):
        g = rng.normal(0, 1, p.size).astype(np.float32)
        net.adam_step(g, 1e-3)
        p, m1, m2 = mlp_ref.adam(p, g.astype(np.float64), m1, m2, step, 1e-3)
    np.testing.assert_allclose(net.get_params(), p, rtol=1e-5, atol=1e-6)


def test_training_fits_a_smooth_target():
    rng = np.random.default_rng(9)
    X = rng.normal(0, 1, (1024, DIMS[0])).astype(np.float32)
    W = rng.normal(0, 0.3, (DIMS[0], DIMS[3]))
    T = np.tanh(X @ W).astype(np.float32)
    net = hip.MLP(*DIMS, seed=11)
    losses = net.train(X, T, batch=128, epochs=60, lr=3e-3)
    print("epoch losses: first %.4f last %.4f" % (losses[0], losses[-1]))
    assert losses[-1] < 0.1 * losses[0]
    # the device parameters still reproduce the oracle's forward pass
    p = net.get_params().astype(np.float64)
    _, acts = mlp_ref.forward(p, DIMS, X[:32].astype(np.float64))
    np.testing.assert_allclose(net.predict(X[:32]), acts[-1], rtol=1e-3, atol=1e-3)


def test_checkpoint_resume_continues_training_exactly():
    """Weights + Adam moments + step saved mid-training and restored into a
    fresh network continue exactly as the uninterrupted run (same kernels,
    same order, so the same bits)."""
    rng = np.random.default_rng(13)
    X = rng.normal(0, 1, (512, DIMS[0])).astype(np.float32)
    T = np.tanh(X @ rng.normal(0, 0.3, (DIMS[0], DIMS[3]))).astype(np.float32)
    a = hip.MLP(*DIMS, seed=5)
    a.train(X, T, batch=64, epochs=3, lr=2e-3)
    params, (m1, m2, step) = a.get_params(), a.get_adam_state()
    assert step > 0 and np.abs(m2).sum() > 0
    a.train(X, T, batch=64, epochs=2, lr=2e-3)
    b = hip.MLP(*DIMS, seed=99)
    b.set_params(params)
    b.set_adam_state(m1, m2, step)
    b.train(X, T, batch=64, epochs=2, lr=2e-3)
    np.testing.assert_array_equal(b.get_params(), a.get_params())
    assert b.get_adam_state()[2] == a.get_adam_state()[2]
