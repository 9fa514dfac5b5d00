"""cp_lstsq_grad on the GPU (reconstruction-free: fused MTTKRP + the
factor-times-accum launch) against the oracle's residual form."""
import numpy as np
import pytest
import torch

from oracle import tensorly_np as T

pytestmark = pytest.mark.gpu


def _case(seed, shape, R, weights=True):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal(shape)
    w = rng.random(R) + 0.5 if weights else None
    A = [rng.standard_normal((s, R)) for s in shape]
    return X, w, A


@pytest.mark.parametrize("shape,R", [((12, 15, 17), 4), ((6, 7, 5, 8), 3), ((40, 33, 64), 32)])
@pytest.mark.parametrize("weights", [True, False])
def test_cp_lstsq_grad_float64_vs_oracle(cuda, shape, R, weights):
    import tensorly_amd as tla
    X, w, A = _case(len(shape) + R, shape, R, weights)
    cp = (None if w is None else torch.tensor(w, device=cuda), [torch.tensor(a, device=cuda) for a in A])
    grad, loss = tla.cp_lstsq_grad(cp, torch.tensor(X, device=cuda), return_loss=True)
    (_, ref), loss_r = T.cp_lstsq_grad((np.ones(R) if w is None else w, A), X, return_loss=True)
    assert grad.weights is None or bool(torch.all(grad.weights == 1))
    for g, r in zip(grad.factors, ref):
        np.testing.assert_allclose(g.cpu().numpy(), r, rtol=1e-9, atol=1e-9 * np.abs(r).max())
    np.testing.assert_allclose(float(loss), loss_r, rtol=1e-10)


def test_cp_lstsq_grad_float32_tracks_oracle(cuda):
    import tensorly_amd as tla
    X, w, A = _case(3, (32, 24, 48), 8)
    cp = (torch.tensor(w, dtype=torch.float32, device=cuda),
          [torch.tensor(a, dtype=torch.float32, device=cuda) for a in A])
    grad = tla.cp_lstsq_grad(cp, torch.tensor(X, dtype=torch.float32, device=cuda))
    _, ref = T.cp_lstsq_grad((w.astype(np.float32).astype(np.float64),
                              [a.astype(np.float32).astype(np.float64) for a in A]),
                             X.astype(np.float32).astype(np.float64))
    for g, r in zip(grad.factors, ref):
        assert np.linalg.norm(g.cpu().numpy() - r) / np.linalg.norm(r) < 1e-5


def test_cp_lstsq_grad_is_zero_at_an_exact_fit(cuda):
    import tensorly_amd as tla
    rng = np.random.default_rng(9)
    A = [rng.standard_normal((s, 3)) for s in (10, 11, 12)]
    X = T.cp_to_tensor((np.ones(3), A))
    grad, loss = tla.cp_lstsq_grad((None, [torch.tensor(a, device=cuda) for a in A]), torch.tensor(X, device=cuda),
                                   return_loss=True)
    for g, a in zip(grad.factors, A):
        assert float(g.abs().max()) < 1e-10 * np.abs(X).max() * a.shape[0]
    assert abs(float(loss)) < 1e-10 * np.sum(X ** 2)


def test_cp_lstsq_grad_rejects_mask(cuda):
    import tensorly_amd as tla
    X = torch.rand((4, 5, 6), dtype=torch.float64, device=cuda)
    with pytest.raises(NotImplementedError):
        tla.cp_lstsq_grad((None, [torch.rand((s, 2), dtype=torch.float64, device=cuda) for s in (4, 5, 6)]), X,
                          mask=torch.ones_like(X))
