"""Oracle functions pinned by brute force, closed forms and LAPACK
identities (no reference table exists for them)."""
import itertools

import numpy as np
import pytest

from oracle import tensorly_np as T


def _brute_mttkrp(X, weights, factors, mode):
    shape = X.shape
    R = factors[0].shape[1]
    out = np.zeros((shape[mode], R))
    for idx in itertools.product(*[range(s) for s in shape]):
        for r in range(R):
            p = X[idx] * (weights[r] if weights is not None else 1.0)
            for m, i in enumerate(idx):
                if m != mode:
                    p *= factors[m][i, r]
            out[idx[mode], r] += p
    return out


@pytest.mark.parametrize("shape", [(3, 4, 5), (2, 3, 2, 3)])
def test_mttkrp_forms_match_brute_force(shape):
    rng = np.random.default_rng(7)
    X = rng.standard_normal(shape)
    R = 3
    factors = [rng.standard_normal((s, R)) for s in shape]
    w = rng.random(R) + 0.5
    for mode in range(len(shape)):
        ref = _brute_mttkrp(X, w, factors, mode)
        np.testing.assert_allclose(T.unfolding_dot_khatri_rao(X, (w, factors), mode), ref, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(T.unfolding_dot_khatri_rao_memory(X, (w, factors), mode), ref, rtol=1e-12,
                                   atol=1e-12)
        np.testing.assert_allclose(T.unfolding_dot_khatri_rao_einsum(X, (w, factors), mode), ref, rtol=1e-12,
                                   atol=1e-12)


def test_planted_mttkrp_mistake_is_caught():
    """Skipping the wrong factor (mode+1) or dropping the weights is caught
    by the brute-force check."""
    rng = np.random.default_rng(8)
    X = rng.standard_normal((3, 4, 5))
    factors = [rng.standard_normal((s, 2)) for s in X.shape]
    w = np.array([2.0, 0.5])
    ref = _brute_mttkrp(X, w, factors, 0)
    no_w = T.unfolding_dot_khatri_rao(X, (None, factors), 0)
    assert not np.allclose(no_w, ref)


def test_mttkrp_closed_form_of_cp_model():
    """X = [[A_m]] gives MTTKRP(X, B, n) = A_n (*_{m != n} A_m^T B_m)."""
    rng = np.random.default_rng(9)
    shape, R1, R2 = (6, 5, 4), 3, 2
    A = [rng.standard_normal((s, R1)) for s in shape]
    B = [rng.standard_normal((s, R2)) for s in shape]
    X = T.cp_to_tensor((np.ones(R1), A))
    for n in range(3):
        H = np.ones((R1, R2))
        for m in range(3):
            if m != n:
                H = H * (A[m].T @ B[m])
        np.testing.assert_allclose(T.unfolding_dot_khatri_rao(X, (None, B), n), A[n] @ H, rtol=1e-11)


def test_cp_norm_and_normalize():
    rng = np.random.default_rng(10)
    w, f = T.random_cp((8, 5, 6, 4), 25, normalise_factors=True, random_state=np.random.RandomState(10))
    np.testing.assert_allclose(T.cp_norm((w, f)), np.linalg.norm(T.cp_to_tensor((w, f))), rtol=1e-12)
    w2, f2 = T.cp_normalize((None, [rng.standard_normal((s, 4)) for s in (3, 4, 5)]))
    for x in f2:
        np.testing.assert_allclose(np.linalg.norm(x, axis=0), np.ones(4), rtol=1e-12)


def test_error_calc_matches_brute_force():
    rng = np.random.default_rng(11)
    X = rng.standard_normal((5, 6, 7))
    w = rng.random(3) + 0.1
    f = [rng.standard_normal((s, 3)) for s in X.shape]
    brute = np.linalg.norm(X - T.cp_to_tensor((w, f)))
    np.testing.assert_allclose(T.error_calc(X, np.linalg.norm(X), w, f), brute, rtol=1e-12)
    m = T.unfolding_dot_khatri_rao(X, (w, f), 2)
    # the fast branch needs the MTTKRP of the last mode with the *current* last factor
    np.testing.assert_allclose(T.error_calc(X, np.linalg.norm(X), w, f, mttkrp=m), brute, rtol=1e-10)


def test_svd_flip_rule():
    U = np.array([[0.1, -0.9], [-0.8, 0.2], [0.3, 0.9]])
    V = np.eye(2)
    U2, V2 = T.svd_flip(U, V)
    # column 0: max |.| is -0.8 -> flipped; column 1: |-0.9| == |0.9|, first wins (-0.9) -> flipped
    np.testing.assert_array_equal(U2, -U)
    np.testing.assert_array_equal(V2, -V)


def test_truncated_and_symeig_svd():
    rng = np.random.default_rng(12)
    A = rng.standard_normal((7, 11))
    U, S, V = T.truncated_svd(A, n_eigenvecs=4)
    assert U.shape == (7, 4) and S.shape == (4,) and V.shape == (4, 11)
    np.testing.assert_allclose(U.T @ U, np.eye(4), atol=1e-12)
    np.testing.assert_allclose(U.T @ A, S[:, None] * V, atol=1e-10)
    U2, S2, V2 = T.symeig_svd(A, n_eigenvecs=4)
    np.testing.assert_allclose(S2, S, rtol=1e-10)
    np.testing.assert_allclose(np.abs(U2.T @ U), np.eye(4), atol=1e-8)
    U3, _, _ = T.svd_interface(A, n_eigenvecs=4)
    idx = np.argmax(np.abs(U3), axis=0)
    assert np.all(U3[idx, np.arange(4)] > 0)


def test_validate_ranks():
    assert T.validate_cp_rank((4, 5, 6), 3) == 3
    assert T.validate_cp_rank((4, 5, 6), "same") == int(np.round(120 / 15))
    assert T.validate_tucker_rank((4, 5, 6), 2) == [2, 2, 2]
    r = T.validate_tucker_rank((10, 10, 10), 0.5)
    assert len(r) == 3 and all(1 <= x <= 10 for x in r)
    with pytest.raises(ValueError):
        T.validate_cp_tensor((np.ones(3), [np.ones((4, 3)), np.ones((5, 2))]))


def test_random_cp_draw_order():
    """init='random' is one random_sample((I_m, R)) per mode in order."""
    w, f = T.random_cp((4, 5, 6), 3, normalise_factors=False, random_state=0)
    rng = np.random.RandomState(0)
    for s, x in zip((4, 5, 6), f):
        np.testing.assert_array_equal(x, rng.random_sample((s, 3)))
    np.testing.assert_array_equal(w, np.ones(3))


def test_cp_lstsq_grad_matches_finite_differences():
    """The gradient of 0.5 ||X - [[w; A]]||^2 against central differences of
    the loss for every factor entry (weights not differentiated)."""
    rng = np.random.default_rng(11)
    shape, R = (3, 4, 5), 2
    X = rng.standard_normal(shape)
    w = rng.random(R) + 0.5
    A = [rng.standard_normal((s, R)) for s in shape]
    (_, grads), loss = T.cp_lstsq_grad((w, A), X, return_loss=True)
    np.testing.assert_allclose(loss, 0.5 * np.sum((X - T.cp_to_tensor((w, A))) ** 2), rtol=1e-13)
    h = 1e-6
    for n in range(3):
        fd = np.zeros_like(A[n])
        for i in range(A[n].shape[0]):
            for r in range(R):
                Ap = [a.copy() for a in A]
                Am = [a.copy() for a in A]
                Ap[n][i, r] += h
                Am[n][i, r] -= h
                fd[i, r] = (0.5 * np.sum((X - T.cp_to_tensor((w, Ap))) ** 2)
                            - 0.5 * np.sum((X - T.cp_to_tensor((w, Am))) ** 2)) / (2 * h)
        np.testing.assert_allclose(grads[n], fd, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("shape", [(7, 8, 9), (4, 5, 3, 6)])
def test_coo_mttkrp_matches_the_densified_tensor(shape):
    rng = np.random.default_rng(21)
    nnz = 60
    coords = np.stack([rng.integers(0, s, nnz) for s in shape])
    coords[:, 1] = coords[:, 0]  # one repeated coordinate: its values add up
    vals = rng.standard_normal(nnz)
    dense = np.zeros(shape)
    np.add.at(dense, tuple(coords), vals)
    R = 3
    w = rng.random(R) + 0.5
    factors = [rng.standard_normal((s, R)) for s in shape]
    for mode in range(len(shape)):
        np.testing.assert_allclose(T.unfolding_dot_khatri_rao_coo(coords, vals, shape, (w, factors), mode),
                                   _brute_mttkrp(dense, w, factors, mode), rtol=1e-12, atol=1e-12)


def test_order2_mttkrp_drops_the_weights():
    """khatri_rao of one matrix returns it as is (``_khatri_rao.py:67-69``),
    so a matrix's MTTKRP is X @ B whatever the weights; the COO restatement
    follows the dense one there."""
    rng = np.random.default_rng(22)
    shape = (6, 9)
    coords = np.stack([rng.integers(0, s, 30) for s in shape])
    vals = rng.standard_normal(30)
    dense = np.zeros(shape)
    np.add.at(dense, tuple(coords), vals)
    w = rng.random(4) + 0.5
    A, B = rng.standard_normal((6, 4)), rng.standard_normal((9, 4))
    assert T.khatri_rao([B], weights=w) is B
    np.testing.assert_allclose(T.unfolding_dot_khatri_rao(dense, (w, [A, B]), 0), dense @ B, rtol=1e-13)
    np.testing.assert_allclose(T.unfolding_dot_khatri_rao(dense, (w, [A, B]), 1), dense.T @ A, rtol=1e-13)
    for mode in range(2):
        np.testing.assert_allclose(T.unfolding_dot_khatri_rao_coo(coords, vals, shape, (w, [A, B]), mode),
                                   T.unfolding_dot_khatri_rao(dense, (w, [A, B]), mode), rtol=1e-12, atol=1e-12)
