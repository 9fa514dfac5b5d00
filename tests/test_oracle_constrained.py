"""The oracle's proximal operators, ADMM and constrained_parafac, and its
non-negative CP, orthogonalised and masked parafac, against outputs recorded
from the reference implementation (tests/golden/make_more_trajectories.py);
plus the properties that pin the proximal operators independently (each is
the argmin it claims to be, checked by brute force or closed form)."""
import os

import numpy as np
import pytest

from oracle import tensorly_np as T

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CONSTRAINTS = {
    "non_negative": True, "l1_reg": 0.05, "l2_reg": 0.5, "l2_square_reg": 0.3, "unimodality": True,
    "normalize": True, "simplex": 3.0, "normalized_sparsity": 14, "soft_sparsity": 2.0, "smoothness": 0.4,
    "monotonicity": True, "hard_sparsity": 17,
}


@pytest.fixture(scope="module")
def more():
    return dict(np.load(os.path.join(GOLDEN, "more_trajectories.npz")))


@pytest.mark.parametrize("name", sorted(CONSTRAINTS))
@pytest.mark.parametrize("key", ["signed", "positive", "negative"])
def test_prox_matches_reference(more, name, key):
    ref_key = f"prox_{name}_{key}"
    if ref_key not in more:
        pytest.skip("the reference raised on this input")
    got = T.proximal_operator(more[f"prox_in_{key}"].copy(), n_const=1, order=0, **{name: CONSTRAINTS[name]})
    np.testing.assert_allclose(got, more[ref_key], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name", sorted(CONSTRAINTS))
def test_admm_matches_reference(more, name):
    x, xs, d = T.admm(more["admm_UtM"], more["admm_UtU"], more["admm_x0"].copy(),
                      np.zeros_like(more["admm_x0"]), n_iter_max=25, n_const=1, order=0, tol=1e-6,
                      **{name: CONSTRAINTS[name]})
    np.testing.assert_allclose(x, more[f"admm_{name}_x"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(xs, more[f"admm_{name}_xsplit"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(d, more[f"admm_{name}_dual"], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("name", sorted(CONSTRAINTS))
@pytest.mark.parametrize("init", ["random", "svd"])
def test_constrained_parafac_matches_reference(more, name, init):
    key = f"cp_{name}_{init}"
    if f"{key}_errors" not in more:
        with pytest.raises(np.linalg.LinAlgError):
            np.random.seed(43)
            T.constrained_parafac(more["cx"], 4, n_iter_max=12, init=init, tol_outer=1e-10, **{name: CONSTRAINTS[name]})
        return
    np.random.seed(43)
    (w, f), errs = T.constrained_parafac(more["cx"], 4, n_iter_max=12, init=init, return_errors=True,
                                         tol_outer=1e-10, **{name: CONSTRAINTS[name]})
    np.testing.assert_allclose(errs, more[f"{key}_errors"], rtol=1e-9)
    for m in range(3):
        np.testing.assert_allclose(f[m], more[f"{key}_factor{m}"], rtol=1e-7, atol=1e-9)


def test_constrained_parafac_mixed_modes(more):
    np.random.seed(44)
    (w, f), errs = T.constrained_parafac(more["cx"], 4, n_iter_max=10, init="random", return_errors=True,
                                         non_negative={0: True}, l1_reg={1: 0.02}, n_iter_max_inner=5)
    np.testing.assert_allclose(errs, more["cp_mixed_errors"], rtol=1e-9)
    for m in range(3):
        np.testing.assert_allclose(f[m], more[f"cp_mixed_factor{m}"], rtol=1e-7, atol=1e-9)


def test_two_constraints_on_one_mode_raise():
    with pytest.raises(ValueError, match="two constraints"):
        T.validate_constraints(n_const=3, non_negative=True, l1_reg={0: 0.1})
    with pytest.raises(ValueError, match="two constraints"):
        T.validate_constraints(n_const=3, non_negative={1: True}, l1_reg=[0, 0.1, 0])


def test_nn_mu_and_hals_match_reference(more):
    (w, f), errs = T.non_negative_parafac(more["nx"], 4, n_iter_max=30, init="random", random_state=3, tol=0,
                                          return_errors=True)
    np.testing.assert_allclose(errs, more["nn_mu_errors"], rtol=1e-9)
    for m in range(3):
        np.testing.assert_allclose(f[m], more[f"nn_mu_factor{m}"], rtol=1e-7, atol=1e-10)
    (w, f), errs = T.non_negative_parafac_hals(more["nx"], 4, n_iter_max=20, init="random", random_state=3, tol=0,
                                               return_errors=True)
    np.testing.assert_allclose(errs, more["nn_hals_errors"], rtol=1e-9)
    for m in range(3):
        np.testing.assert_allclose(f[m], more[f"nn_hals_factor{m}"], rtol=1e-7, atol=1e-10)


@pytest.mark.parametrize("k", [True, 4])
def test_parafac_orthogonalise_matches_reference(more, k):
    (w, f), errs = T.parafac(more["ox"], 3, n_iter_max=8, init="random", random_state=5, tol=0, orthogonalise=k,
                             return_errors=True)
    np.testing.assert_allclose(errs, more[f"orth_{int(k)}_errors"], rtol=1e-9)
    for m in range(3):
        np.testing.assert_allclose(f[m], more[f"orth_{int(k)}_factor{m}"], rtol=1e-7, atol=1e-10)


# -- properties that pin the proximal operators without the recording -------
def _brute_argmin(obj, x0, steps=4000):
    """Minimise a convex objective by a fine coordinate search (tiny cases)."""
    from scipy.optimize import minimize
    res = minimize(obj, x0, method="Nelder-Mead", options={"xatol": 1e-10, "fatol": 1e-14, "maxiter": steps})
    return res.x


def test_soft_thresholding_is_the_l1_prox():
    rng = np.random.default_rng(1)
    v = rng.standard_normal(5)
    lam = 0.3
    x = T.soft_thresholding(v, lam)
    y = _brute_argmin(lambda z: lam * np.sum(np.abs(z)) + 0.5 * np.sum((z - v) ** 2), v)
    np.testing.assert_allclose(x, y, atol=1e-5)


def test_simplex_prox_is_the_euclidean_projection():
    rng = np.random.default_rng(2)
    v = rng.standard_normal((6, 3))
    x = T.simplex_prox(v, 2.0)
    assert np.all(x >= 0)
    np.testing.assert_allclose(x.sum(axis=0), 2.0, rtol=1e-12)
    # KKT of min ||x - v||^2 on the simplex: x = max(v - tau, 0) with one tau
    for c in range(3):
        s = x[:, c] > 0
        tau = (v[s, c] - x[s, c])
        np.testing.assert_allclose(tau, tau[0], atol=1e-12)
        assert np.all(v[~s, c] <= tau[0] + 1e-12)


def test_smoothness_prox_solves_the_banded_system():
    rng = np.random.default_rng(3)
    v = rng.standard_normal((7, 2))
    lam = 0.7
    x = T.smoothness_prox(v, lam)
    # (2 lam + 1) x_i - lam (x_{i-1} + x_{i+1}) = v_i, zero outside the range
    xp = np.vstack([np.zeros((1, 2)), x, np.zeros((1, 2))])
    np.testing.assert_allclose((2 * lam + 1) * xp[1:-1] - lam * (xp[:-2] + xp[2:]), v, atol=1e-12)


def test_monotonicity_prox_output_is_nondecreasing():
    rng = np.random.default_rng(4)
    v = rng.standard_normal((12, 4))
    x = T.monotonicity_prox(v)
    assert np.all(np.diff(x, axis=0) >= -1e-15)
    xd = T.monotonicity_prox(v, decreasing=True)
    assert np.all(np.diff(xd, axis=0) <= 1e-15)
    # an already-sorted non-negative column is a fixed point
    s = np.sort(np.abs(v), axis=0)
    np.testing.assert_allclose(T.monotonicity_prox(s), s, rtol=1e-14)


def test_hard_thresholding_keeps_the_largest():
    v = np.array([[0.5, -3.0], [2.0, 0.1], [-1.0, 4.0]])
    np.testing.assert_array_equal(T.hard_thresholding(v, 3), [[0.0, -3.0], [2.0, 0.0], [0.0, 4.0]])


def test_planted_prox_mistake_is_caught(more):
    """A simplex projection that forgets the parameter (projects on the unit
    simplex) fails the recorded output."""
    x = more["prox_in_signed"]
    wrong = T.simplex_prox(x, 1.0)
    assert not np.allclose(wrong, more["prox_simplex_signed"], atol=1e-6)


@pytest.mark.parametrize("init", ["random", "svd"])
def test_parafac_mask_matches_reference(more, init):
    kw = dict(init="random", random_state=6) if init == "random" else dict(init="svd", svd_mask_repeats=3)
    (w, f), errs = T.parafac(more["mx"], 3, n_iter_max=10, tol=0, mask=more["mmask"], return_errors=True, **kw)
    np.testing.assert_allclose(errs, more[f"mask_{init}_errors"], rtol=1e-9)
    for m in range(3):
        np.testing.assert_allclose(f[m], more[f"mask_{init}_factor{m}"], rtol=1e-7, atol=1e-10)
