"""MTTKRP of a sparse COO tensor on the GPU (tla_mttkrp_coo) against the
oracle and against the dense MTTKRP of the densified tensor."""
import numpy as np
import pytest
import torch

from oracle import tensorly_np as T

pytestmark = pytest.mark.gpu


def _coo(seed, shape, nnz, dup=True):
    rng = np.random.default_rng(seed)
    coords = np.stack([rng.integers(0, s, nnz) for s in shape]).astype(np.int64)
    if dup and nnz > 2:
        coords[:, 1] = coords[:, 0]
    vals = rng.standard_normal(nnz)
    return coords, vals


@pytest.mark.parametrize("shape,nnz,R", [((30, 20, 40), 500, 5), ((12, 9, 7, 11), 800, 16), ((50, 60, 70), 1, 3),
                                         ((64, 64, 64), 20000, 64)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_coo_mttkrp_vs_oracle_and_dense(cuda, shape, nnz, R, dtype):
    import tensorly_amd as tla
    coords, vals = _coo(len(shape) + nnz, shape, nnz)
    rng = np.random.default_rng(R)
    w = rng.random(R) + 0.5
    factors = [rng.standard_normal((s, R)) for s in shape]
    c = torch.tensor(coords, device=cuda)
    v = torch.tensor(vals, dtype=dtype, device=cuda)
    wt = torch.tensor(w, dtype=dtype, device=cuda)
    ft = [torch.tensor(f, dtype=dtype, device=cuda) for f in factors]
    dense = np.zeros(shape)
    np.add.at(dense, tuple(coords), vals)
    for mode in range(len(shape)):
        got = tla.tenalg.unfolding_dot_khatri_rao_coo(c, v, shape, (wt, ft), mode)
        ref = T.unfolding_dot_khatri_rao_coo(coords, vals, shape, (w, factors), mode)
        tol = 1e-12 if dtype == torch.float64 else 2e-5
        np.testing.assert_allclose(got.double().cpu().numpy(), ref, rtol=tol, atol=tol * np.abs(ref).max())
        if dtype == torch.float64:
            dense_gpu = tla.unfolding_dot_khatri_rao(torch.tensor(dense, device=cuda), (wt, ft), mode)
            np.testing.assert_allclose(got.cpu().numpy(), dense_gpu.cpu().numpy(), rtol=1e-10,
                                       atol=1e-10 * np.abs(ref).max())
        again = tla.tenalg.unfolding_dot_khatri_rao_coo(c, v, shape, (wt, ft), mode)
        assert torch.equal(got, again)


def test_coo_mttkrp_is_sparse_safe(cuda):
    """The reference's sparse-safety case: shape (1000, 1000, 1000, 10),
    rank 5, here with 2M nonzeros; the dense tensor (10^10 entries) is never
    formed."""
    import tensorly_amd as tla
    shape = (1000, 1000, 1000, 10)
    coords, vals = _coo(3, shape, 2_000_000, dup=False)
    rng = np.random.default_rng(4)
    factors = [rng.random((s, 5)) for s in shape]
    c = torch.tensor(coords, device=cuda)
    v = torch.tensor(vals, device=cuda)
    ft = [torch.tensor(f, device=cuda) for f in factors]
    for mode in range(4):
        got = tla.tenalg.unfolding_dot_khatri_rao_coo(c, v, shape, (None, ft), mode)
        ref = T.unfolding_dot_khatri_rao_coo(coords, vals, shape, (None, factors), mode)
        np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())


def test_coo_mttkrp_rejects_out_of_range_coordinates(cuda):
    import tensorly_amd as tla
    shape = (5, 6, 7)
    coords = torch.tensor([[0, 4], [0, 6], [0, 1]], dtype=torch.int64, device=cuda)  # 6 >= shape[1]
    v = torch.ones(2, dtype=torch.float64, device=cuda)
    ft = [torch.ones((s, 2), dtype=torch.float64, device=cuda) for s in shape]
    with pytest.raises(ValueError, match="out of the tensor"):
        tla.tenalg.unfolding_dot_khatri_rao_coo(coords, v, shape, (None, ft), 0)
