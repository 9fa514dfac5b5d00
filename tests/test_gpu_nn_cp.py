"""non_negative_parafac (multiplicative updates) on the GPU against the
oracle, and the tla_nn_mu_update kernel against the update formula."""
import numpy as np
import pytest
import torch

from oracle import tensorly_np as T

pytestmark = pytest.mark.gpu


def _data(seed, shape, rank, noise=0.02):
    rng = np.random.RandomState(seed)
    A = [rng.random_sample((s, rank)) + 0.1 for s in shape]
    X = T.cp_to_tensor((np.ones(rank), A))
    return X + noise * rng.random_sample(shape) * X.mean()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("rows,rank", [(1, 1), (37, 5), (300, 16), (129, 64)])
def test_mu_update_kernel_matches_formula(cuda, dtype, rows, rank):
    from tensorly_amd.cp_tensor import grams64
    from tensorly_amd.decomposition._nn_cp import _mu_update
    rng = np.random.RandomState(rows + rank)
    shape = (rows, 20, 9)
    factors = [torch.tensor(rng.random_sample((s, rank)), dtype=dtype, device=cuda) for s in shape]
    w = torch.tensor(rng.random_sample(rank) + 0.5, dtype=dtype, device=cuda)
    mttkrp = torch.tensor(rng.standard_normal((rows, rank)), dtype=dtype, device=cuda)  # negatives exercise the clip
    A0 = factors[0].double().cpu().numpy()
    grams = grams64(factors)
    eps = float(torch.finfo(dtype).eps)
    _mu_update(factors, mttkrp, grams, w, 0, eps)
    Fn = [f.double().cpu().numpy() for f in factors]
    acc = (Fn[1].T @ Fn[1]) * (Fn[2].T @ Fn[2])
    wn = w.double().cpu().numpy()
    acc = wn[:, None] * acc * wn[None, :]
    ref = A0 * np.clip(mttkrp.double().cpu().numpy(), eps, None) / np.clip(A0 @ acc, eps, None)
    tol = 1e-6 if dtype == torch.float32 else 1e-13
    np.testing.assert_allclose(factors[0].double().cpu().numpy(), ref, rtol=tol, atol=0)


def test_mu_update_rejects_bad_rank(cuda):
    from tensorly_amd.decomposition._nn_cp import _mu_update
    factors = [torch.rand((4, 65), device=cuda) for _ in range(3)]
    grams = [torch.zeros((65, 65), dtype=torch.float64, device=cuda) for _ in range(3)]
    with pytest.raises(ValueError, match="rank 65"):
        _mu_update(factors, torch.rand((4, 65), device=cuda), grams, torch.ones(65, device=cuda), 0, 1e-7)


@pytest.mark.parametrize("init", ["random", "svd"])
@pytest.mark.parametrize("shape,rank", [((16, 20, 24), 4), ((6, 8, 5, 7), 3)])
def test_non_negative_parafac_float64_vs_oracle(cuda, init, shape, rank):
    import tensorly_amd as tla
    X = _data(1, shape, rank)
    cp, errs = tla.non_negative_parafac(torch.tensor(X, device=cuda), rank, n_iter_max=40, init=init,
                                        random_state=3, tol=1e-300, return_errors=True)
    (w_r, f_r), errs_r = T.non_negative_parafac(X, rank, n_iter_max=40, init=init, random_state=3, tol=1e-300,
                                                return_errors=True)
    assert len(errs) == len(errs_r) == 40
    # the GPU SVD (FP64 Gram eigensolve) and LAPACK agree to ~1e-12 on the
    # leading vectors; the random init is bit-identical
    np.testing.assert_allclose(errs, errs_r, rtol=1e-9 if init == "random" else 1e-6)
    for f, fr in zip(cp.factors, f_r):
        a = f.cpu().numpy()
        assert (a >= 0).all()
        np.testing.assert_allclose(a, fr, rtol=1e-6 if init == "random" else 1e-3,
                                   atol=1e-9 if init == "random" else 1e-5)


def test_non_negative_parafac_float32_tracks_oracle(cuda):
    import tensorly_amd as tla
    X = _data(2, (24, 32, 40), 5)
    cp, errs = tla.non_negative_parafac(torch.tensor(X, dtype=torch.float32, device=cuda), 5, n_iter_max=60,
                                        init="random", random_state=0, tol=1e-300, return_errors=True)
    _, errs_r = T.non_negative_parafac(X, 5, n_iter_max=60, init="random", random_state=0, tol=1e-300,
                                       return_errors=True)
    # float32 MTTKRP and factors against the FP64 oracle; the fit from Grams
    # carries the float32 rounding of <X, M>
    np.testing.assert_allclose(errs, errs_r, atol=2e-5)
    assert all((f >= 0).all() for f in cp.factors)


@pytest.mark.parametrize("kw", [{"normalize_factors": True}, {"fixed_modes": [1]},
                                {"cvg_criterion": "rec_error", "tol": 1e-4}])
def test_non_negative_parafac_options_vs_oracle(cuda, kw):
    import tensorly_amd as tla
    X = _data(4, (10, 12, 14), 3)
    kw = dict(kw)
    tol = kw.pop("tol", 1e-300)
    cp, errs = tla.non_negative_parafac(torch.tensor(X, device=cuda), 3, n_iter_max=30, init="random",
                                        random_state=5, tol=tol, return_errors=True, **kw)
    (w_r, f_r), errs_r = T.non_negative_parafac(X, 3, n_iter_max=30, init="random", random_state=5, tol=tol,
                                                return_errors=True, **kw)
    assert len(errs) == len(errs_r)
    np.testing.assert_allclose(errs, errs_r, rtol=1e-9)
    np.testing.assert_allclose(cp.weights.cpu().numpy(), w_r, rtol=1e-8)
    if "fixed_modes" in kw:
        init = T.initialize_cp(X, 3, init="random", random_state=5, non_negative=True)[1]
        np.testing.assert_array_equal(cp.factors[1].cpu().numpy(), init[1])


def test_non_negative_parafac_rejects_mask(cuda):
    import tensorly_amd as tla
    X = torch.rand((4, 5, 6), device=cuda, dtype=torch.float64)
    with pytest.raises(NotImplementedError):
        tla.non_negative_parafac(X, 2, mask=torch.ones_like(X))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("rows,rank", [(1, 1), (37, 5), (700, 16), (300, 24), (129, 33), (90, 40), (64, 64)])
@pytest.mark.parametrize("sparsity", [None, 0.05])
def test_hals_nnls_kernel_vs_oracle(cuda, dtype, rows, rank, sparsity):
    from tensorly_amd.cp_tensor import grams64
    from tensorly_amd.decomposition._nn_cp import _hals_update
    rng = np.random.RandomState(rows * 7 + rank)
    shape = (rows, 11, 9)
    factors = [torch.tensor(rng.random_sample((s, rank)), dtype=dtype, device=cuda) for s in shape]
    w = torch.tensor(rng.random_sample(rank) + 0.5, dtype=dtype, device=cuda)
    mttkrp = torch.tensor(rng.standard_normal((rows, rank)) * 3.0, dtype=dtype, device=cuda)
    A0 = factors[0].double().cpu().numpy()
    grams = grams64(factors)
    _hals_update(factors, mttkrp, grams, w, 0, sparsity, 100, 1e-8)
    Fn = [f.double().cpu().numpy() for f in factors]
    wn = w.double().cpu().numpy()
    utu = wn[:, None] * ((Fn[1].T @ Fn[1]) * (Fn[2].T @ Fn[2])) * wn[None, :]
    ref = T.hals_nnls(mttkrp.double().cpu().numpy().T, utu, A0.T, n_iter_max=100, tol=1e-8,
                      sparsity_coefficient=sparsity).T
    got = factors[0].double().cpu().numpy()
    assert (got >= 0).all()
    if dtype == torch.float64:
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    else:
        assert np.linalg.norm(got - ref) <= 1e-4 * max(np.linalg.norm(ref), 1e-30)


@pytest.mark.parametrize("init", ["random", "svd"])
@pytest.mark.parametrize("kw", [{}, {"normalize_factors": True}, {"fixed_modes": [0]}, {"nn_modes": [0, 2]},
                                {"sparsity_coefficients": [0.01, None, 0.02]}])
def test_non_negative_parafac_hals_float64_vs_oracle(cuda, init, kw):
    import tensorly_amd as tla
    X = _data(6, (14, 12, 16), 3)
    cp, errs = tla.non_negative_parafac_hals(torch.tensor(X, device=cuda), 3, n_iter_max=15, init=init,
                                             random_state=4, tol=1e-300, return_errors=True, **kw)
    (w_r, f_r), errs_r = T.non_negative_parafac_hals(X, 3, n_iter_max=15, init=init, random_state=4, tol=1e-300,
                                                     return_errors=True, **kw)
    # tol=1e-300 stops only at an exactly repeated error, which the GPU's and
    # the oracle's last digits can reach one iteration apart
    n = min(len(errs), len(errs_r))
    assert n >= 10 and abs(len(errs) - len(errs_r)) <= 1
    np.testing.assert_allclose(errs[:n], errs_r[:n], rtol=1e-8 if init == "random" else 1e-5)
    for f, fr, m in zip(cp.factors, f_r, range(3)):
        a = f.cpu().numpy()
        if "nn_modes" not in kw or m in kw["nn_modes"]:
            assert (a >= 0).all()
        np.testing.assert_allclose(a, fr, rtol=1e-6 if init == "random" else 1e-3,
                                   atol=1e-8 * np.abs(fr).max() if init == "random" else 1e-4 * np.abs(fr).max())


def test_non_negative_parafac_hals_exact_and_float32(cuda):
    import tensorly_amd as tla
    X = _data(7, (10, 9, 12), 2)
    cp, errs = tla.non_negative_parafac_hals(torch.tensor(X, device=cuda), 2, n_iter_max=5, init="random",
                                             random_state=0, tol=1e-300, exact=True, return_errors=True)
    _, errs_r = T.non_negative_parafac_hals(X, 2, n_iter_max=5, init="random", random_state=0, tol=1e-300,
                                            exact=True, return_errors=True)
    np.testing.assert_allclose(errs, errs_r, rtol=1e-7)
    cp32, errs32 = tla.non_negative_parafac_hals(torch.tensor(X, dtype=torch.float32, device=cuda), 2,
                                                 n_iter_max=20, init="random", random_state=0, tol=1e-300,
                                                 return_errors=True)
    _, errs_r32 = T.non_negative_parafac_hals(X, 2, n_iter_max=20, init="random", random_state=0, tol=1e-300,
                                              return_errors=True)
    np.testing.assert_allclose(errs32, errs_r32, atol=2e-5)
