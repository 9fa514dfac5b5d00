"""numpy oracle of the extrapolation MLP (16 -> 64 -> 64 -> 1, tanh, MSE,
Adam).  TEST INFRASTRUCTURE ONLY.

There is no reference counterpart (the learned step is an extension of the
reference's line search, tensorly/decomposition/_cp.py:448-483).  Parity of
the HIP kernels is pinned by: this restatement's gradient against central
finite differences (float64), the HIP gradient against this gradient, and
Adam steps against the update rule written out below (Kingma & Ba, 2015).
"""
import numpy as np

NF, NH = 16, 64


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NH * NF].reshape(NH, NF); o += NH * NF
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    w3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    H1 = np.tanh(X @ W1.T + b1)
    H2 = np.tanh(H1 @ W2.T + b2)
    return H2 @ w3 + b3, (H1, H2)


def loss(p, X, y):
    yh, _ = forward(p, X)
    return float(np.mean((yh - y) ** 2))


def grad(p, X, y):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    yh, (H1, H2) = forward(p, X)
    B = X.shape[0]
    gy = 2.0 * (yh - y) / B
    dw3 = H2.T @ gy
    db3 = gy.sum()
    d2 = gy[:, None] * w3[None, :] * (1 - H2 ** 2)
    dW2 = d2.T @ H1
    db2 = d2.sum(0)
    d1 = (d2 @ W2) * (1 - H1 ** 2)
    dW1 = d1.T @ X
    db1 = d1.sum(0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
