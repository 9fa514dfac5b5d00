"""The oracle's non-negative CP (multiplicative updates) and NNDSVD pinned by
invariants of the mathematics: the Lee-Seung monotone decrease of the
error, exact factorisations as fixed points, NNDSVD reproducing a
non-negative rank-1 matrix, and non-negativity of every output."""
import numpy as np
import pytest

from oracle import tensorly_np as T


def _positive_model(seed, shape, rank):
    rng = np.random.RandomState(seed)
    A = [rng.random_sample((s, rank)) + 0.1 for s in shape]
    return T.cp_to_tensor((np.ones(rank), A)), A


@pytest.mark.parametrize("init", ["random", "svd"])
@pytest.mark.parametrize("shape", [(9, 11, 13), (5, 6, 4, 7)])
def test_mu_error_never_increases(init, shape):
    """Multiplicative updates do not increase ||X - [[A]]|| (Lee and Seung,
    2001): every per-iteration error is at most the previous one."""
    X, _ = _positive_model(3, shape, 3)
    X = X + 0.05 * np.random.RandomState(4).random_sample(shape)
    _, errs = T.non_negative_parafac(X, 4, n_iter_max=60, init=init, random_state=2, tol=1e-300,
                                     return_errors=True)
    assert len(errs) == 60
    assert np.all(np.diff(errs) <= 1e-12 * errs[0])
    assert errs[-1] < errs[0]


def test_mu_leaves_an_exact_factorisation_unchanged():
    """When X = [[w; A]] exactly, M = A accum for every mode, so
    A * M / (A accum) = A: the exact factors (normalised, so the weights are
    not ones and the w w^T scaling of accum is exercised) are a fixed point."""
    X, A = _positive_model(5, (7, 8, 9), 3)
    w, An = T.cp_normalize((np.ones(3), A))
    assert not np.allclose(w, 1.0)
    (w2, A2), errs = T.non_negative_parafac(X, 3, n_iter_max=3, init=(np.ones(3), [a.copy() for a in A]),
                                            normalize_factors=True, tol=1e-300, return_errors=True)
    rec = T.cp_to_tensor((w2, A2))
    np.testing.assert_allclose(rec, X, rtol=1e-10, atol=1e-12)
    # the error from Grams and the MTTKRP cancels: an exact model reads ~sqrt(eps)
    assert max(errs) < 1e-7
    for a, b in zip(A2, An):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)


def test_mu_update_moves_a_perturbed_factor_towards_the_solution():
    X, A = _positive_model(6, (6, 7, 8), 2)
    B = [a.copy() for a in A]
    B[0] = B[0] * 1.3
    _, errs = T.non_negative_parafac(X, 2, n_iter_max=1, init=(np.ones(2), B), tol=1e-300,
                                     return_errors=True)
    e0 = T.norm(X - T.cp_to_tensor((np.ones(2), B))) / T.norm(X)
    assert errs[0] < e0


def test_nndsvd_reproduces_a_nonnegative_rank_one_matrix():
    rng = np.random.RandomState(7)
    u, v = rng.random_sample(12) + 0.2, rng.random_sample(30) + 0.2
    M = np.outer(u, v)
    W, S, H = T.svd_interface(M, n_eigenvecs=1, non_negative="nndsvd")
    # W[:, 0] H[0] = S0 |U0| |V0| = u v^T
    np.testing.assert_allclose(np.outer(W[:, 0], H[0]), M, rtol=1e-10)


@pytest.mark.parametrize("nntype", [True, "nndsvd", "nndsvda"])
def test_nndsvd_outputs_are_nonnegative_and_keep_the_leading_triplet(nntype):
    rng = np.random.RandomState(8)
    M = rng.standard_normal((15, 40))
    U0, S0, V0 = T.svd_interface(M, n_eigenvecs=5)
    W, H = T.make_svd_non_negative(M, U0, S0, V0, nntype)
    assert W.shape == U0.shape and H.shape == V0.shape
    assert (W >= 0).all() and (H >= 0).all()
    # column 0: sqrt(S0) |U0| (entries under eps replaced by the fill rule)
    eps = np.finfo(M.dtype).eps
    w0 = np.sqrt(S0[0]) * np.abs(U0[:, 0])
    keep = w0 >= eps
    np.testing.assert_allclose(W[keep, 0], w0[keep], rtol=1e-12)
    if nntype in (True, "nndsvda"):
        # every entry that NNDSVD thresholds to zero holds the matrix mean instead
        raw = T.make_svd_non_negative(M, U0, S0, V0, "nndsvd")[0]
        small = raw == 0.0
        assert small.any()
        np.testing.assert_allclose(W[small], M.mean())


def test_nndsvd_column_takes_the_dominant_sign_part():
    """For a column pair whose positive parts carry more mass than the
    negative parts, NNDSVD keeps the positive parts scaled by
    sqrt(S_j * |x_p| |y_p|)."""
    U = np.array([[0.8, 0.6], [0.6, -0.1], [0.0, 0.3]])
    V = np.array([[0.5, 0.5, 0.5, 0.5], [0.7, 0.2, -0.1, 0.0]])
    S = np.array([3.0, 2.0])
    W, H = T.make_svd_non_negative(np.ones((3, 4)), U, S, V, "nndsvd")
    x_p = np.array([0.6, 0.0, 0.3])
    y_p = np.array([0.7, 0.2, 0.0, 0.0])
    sigma = np.linalg.norm(x_p) * np.linalg.norm(y_p)
    lbd = np.sqrt(2.0 * sigma)
    eps = np.finfo(np.float64).eps
    np.testing.assert_allclose(W[:, 1], np.clip(lbd * x_p / np.linalg.norm(x_p) - eps, 0, None), rtol=1e-12)
    np.testing.assert_allclose(H[1], np.clip(lbd * y_p / np.linalg.norm(y_p) - eps, 0, None), rtol=1e-12)


def test_hals_nnls_exact_meets_the_kkt_conditions():
    """min_{V >= 0} ||M - U V||^2 at the solution: V >= 0, the gradient
    UtU V - UtM vanishes where V > 0 and is >= 0 where V = 0."""
    rng = np.random.RandomState(12)
    U = rng.random_sample((30, 5))
    M = rng.standard_normal((30, 8))
    UtU, UtM = U.T @ U, U.T @ M
    V = T.hals_nnls(UtM, UtU, rng.random_sample((5, 8)), exact=True)
    grad = UtU @ V - UtM
    scale = np.abs(UtM).max()
    assert (V >= 0).all() and (V == 0).any()
    # exact mode stops when an iteration's squared update is 1e-16 of the
    # first one's, which leaves gradients of ~1e-8 relative
    assert np.abs(grad[V > 0]).max() < 1e-6 * scale
    assert grad[V == 0].min() > -1e-6 * scale


@pytest.mark.parametrize("init", ["random", "svd"])
def test_hals_error_never_increases(init):
    """Each row update of HALS minimises the objective exactly over that row
    (within the constraint), so the error never increases."""
    X, _ = _positive_model(13, (8, 9, 10), 3)
    X = X + 0.05 * np.random.RandomState(14).random_sample(X.shape)
    _, errs = T.non_negative_parafac_hals(X, 4, n_iter_max=40, init=init, random_state=2, tol=1e-300,
                                          return_errors=True)
    assert len(errs) == 40
    assert np.all(np.diff(errs) <= 1e-12 * errs[0])


def test_hals_leaves_an_exact_factorisation_unchanged():
    X, A = _positive_model(15, (6, 7, 8), 3)
    (w, A2), errs = T.non_negative_parafac_hals(X, 3, n_iter_max=3, init=(np.ones(3), [a.copy() for a in A]),
                                               tol=1e-300, return_errors=True)
    for a, b in zip(A2, A):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)
    assert max(errs) < 1e-7
