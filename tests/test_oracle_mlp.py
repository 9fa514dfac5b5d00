"""The MLP oracle's backward pass against central finite differences, and
the Adam rule against its closed form on a 1-D quadratic."""
import numpy as np

from oracle import mlp_np as M


def _params(seed=0):
    rng = np.random.RandomState(seed)
    n = M.NH * M.NF + M.NH + M.NH * M.NH + M.NH + M.NH + 1
    return rng.standard_normal(n) * 0.2


def test_param_count():
    assert _params().size == 5313


def test_gradient_matches_finite_differences():
    rng = np.random.RandomState(1)
    p = _params(1)
    X = rng.standard_normal((12, M.NF))
    y = rng.standard_normal(12)
    g = M.grad(p, X, y)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss(p + e, X, y) - M.loss(p - e, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_planted_backward_mistake_is_caught():
    """Dropping the tanh derivative of layer 1 changes the W1 gradient."""
    rng = np.random.RandomState(2)
    p = _params(2)
    X = rng.standard_normal((8, M.NF))
    y = rng.standard_normal(8)
    W1, b1, W2, b2, w3, b3 = M.unpack(p)
    yh, (H1, H2) = M.forward(p, X)
    gy = 2.0 * (yh - y) / 8
    d2 = gy[:, None] * w3[None, :] * (1 - H2 ** 2)
    wrong_dW1 = ((d2 @ W2)).T @ X  # missing (1 - H1^2)
    assert not np.allclose(wrong_dW1.ravel(), M.grad(p, X, y)[: M.NH * M.NF])


def test_adam_first_step_is_signed_lr():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    p2, m, v = M.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(p2, p - 0.01 * np.sign(g), rtol=1e-4)
