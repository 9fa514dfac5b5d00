"""The oracle against the literal tables of the reference test-suite
(tests/golden/known_answers.json), plus planted-mistake checks showing the
tables catch plausible errors (a swapped Khatri-Rao order, a transposed
operand, a dropped weight)."""
import numpy as np
import pytest

from oracle import tensorly_np as T
from conftest import arange_matrix


def test_unfold_tables(known_answers):
    ka = known_answers["unfold_arange24_342"]
    X = np.arange(24).reshape(ka["shape"])
    for mode, expected in enumerate(ka["unfoldings"]):
        np.testing.assert_array_equal(T.unfold(X, mode), np.array(expected))
        np.testing.assert_array_equal(T.fold(np.array(expected), mode, X.shape), X)


def test_fold_roundtrip_4way():
    X = np.random.default_rng(0).random((2, 3, 4, 5))
    for m in range(4):
        np.testing.assert_array_equal(T.fold(T.unfold(X, m), m, X.shape), X)


def test_khatri_rao_tables(known_answers):
    for key in ("kr_classic", "kr_eye"):
        ka = known_answers[key]
        mats = [np.array(m, dtype=float) for m in ka["inputs"]]
        np.testing.assert_array_equal(T.khatri_rao(mats), np.array(ka["expected"]))
    ka = known_answers["kr_arange4"]
    U = [arange_matrix(*spec) for spec in ka["inputs_arange"]]
    np.testing.assert_array_equal(T.khatri_rao(U), np.array(ka["expected"]))
    np.testing.assert_array_equal(T.khatri_rao(U, skip_matrix=1), T.khatri_rao([U[0]] + U[2:]))
    np.testing.assert_array_equal(T.khatri_rao(U, skip_matrix=0), T.khatri_rao(U[1:]))
    np.testing.assert_array_equal(T.khatri_rao([U[0]]), U[0])


def test_khatri_rao_errors():
    with pytest.raises(ValueError):
        T.khatri_rao([np.zeros((3, 4)), np.zeros((3, 4)), np.zeros((3, 2))])
    with pytest.raises(ValueError):
        T.khatri_rao([np.eye(3), np.arange(12).reshape(3, 2, 2)])


def test_planted_kr_order_mistake_is_caught(known_answers):
    """A restatement with the LAST matrix varying slowest fails the table."""
    ka = known_answers["kr_arange4"]
    U = [arange_matrix(*spec) for spec in ka["inputs_arange"]]
    wrong = T.khatri_rao(U[::-1])
    assert not np.array_equal(wrong, np.array(ka["expected"]))


def test_mode_dot_tables(known_answers):
    ka = known_answers["mode_dot_matrix"]
    np.testing.assert_array_equal(T.mode_dot(np.array(ka["X"]), np.array(ka["U"]), ka["mode"]),
                                  np.array(ka["expected"]))
    ka = known_answers["mode_dot_vector"]
    np.testing.assert_array_equal(T.mode_dot(np.array(ka["X"]), np.array(ka["v"]), ka["mode"]),
                                  np.array(ka["expected"]))
    ka = known_answers["mode_dot_vector_arange"]
    X = np.arange(ka["X_arange"][0], dtype=float).reshape(ka["X_arange"][1])
    np.testing.assert_array_equal(T.mode_dot(X, np.arange(ka["v_arange"], dtype=float), ka["mode"]),
                                  np.array(ka["expected"]))
    ka = known_answers["multi_mode_dot_vectors"]
    res = T.multi_mode_dot(np.array(ka["X"]), [np.array(u) for u in ka["U"]], ka["modes"])
    np.testing.assert_array_equal(np.reshape(res, (-1,)), np.array(ka["expected"]))


def test_planted_unfold_order_mistake_is_caught(known_answers):
    """A Fortran-order unfolding (columns in the wrong order) fails the
    unfold table and, folded back with the C-order fold, the mode_dot table."""
    ka = known_answers["unfold_arange24_342"]
    X = np.arange(24).reshape(ka["shape"])
    wrong_unfold = np.reshape(np.moveaxis(X, 1, 0), (4, -1), order="F")
    assert not np.array_equal(wrong_unfold, np.array(ka["unfoldings"][1]))
    kb = known_answers["mode_dot_matrix"]
    Xm, U = np.array(kb["X"]), np.array(kb["U"])
    wrong = T.fold(np.dot(U, np.reshape(Xm, (3, -1), order="F")), 0, (2, 4, 2))
    assert not np.array_equal(wrong, np.array(kb["expected"]))


def test_mode_dot_errors():
    X = np.random.default_rng(1).random((2, 4, 5))
    U = np.random.default_rng(2).random((3, 4))
    with pytest.raises(ValueError):
        T.mode_dot(X, U, 0)
    with pytest.raises(ValueError):
        T.mode_dot(X, U[:, 0], 0)
    with pytest.raises(ValueError):
        T.mode_dot(X, X, 0)


def test_multi_mode_dot_skip_and_vectors():
    rng = np.random.default_rng(3)
    dims = [2, 3, 4, 5]
    X = rng.standard_normal(dims)
    factors = [rng.random((dims[i], dims[i])) for i in range(4)]
    np.testing.assert_array_equal(T.multi_mode_dot(X, factors, skip=1),
                                  T.multi_mode_dot(X, [factors[0]] + factors[2:], modes=[0, 2, 3]))
    shape = (3, 5, 4, 2)
    ones = np.ones(shape)
    vecs = [np.ones(s) / s for s in shape]
    import itertools
    for modes in itertools.permutations(range(4)):
        res = T.multi_mode_dot(ones, [vecs[i] for i in modes], modes=modes)
        assert np.shape(res) == ()
        np.testing.assert_allclose(res, 1.0)


def test_multi_mode_dot_kronecker_identity():
    rng = np.random.default_rng(4)
    dims = [2, 3, 4, 5]
    X = rng.standard_normal(dims)
    factors = [rng.random((dims[i] + 1, dims[i])) for i in range(4)]
    res = T.multi_mode_dot(X, factors)
    for i in range(4):
        others = [factors[j] for j in range(4) if j != i]
        kron = others[0]
        for f in others[1:]:
            kron = np.kron(kron, f)
        np.testing.assert_allclose(T.unfold(res, i), factors[i] @ T.unfold(X, i) @ kron.T, rtol=1e-12)


def test_cp_to_tensor_tables(known_answers):
    ka = known_answers["cp_to_tensor_arange4"]
    U = [arange_matrix(*spec) for spec in ka["inputs_arange"]]
    np.testing.assert_array_equal(T.cp_to_tensor((np.ones(3), U)), np.array(ka["expected"]))
    ka = known_answers["cp_to_tensor_weights"]
    A = np.arange(*ka["A_arange"][:2], dtype=float).reshape(ka["A_arange"][2:])
    B = np.arange(*ka["B_arange"][:2], dtype=float).reshape(ka["B_arange"][2:])
    np.testing.assert_array_equal(T.cp_to_tensor((np.array(ka["weights"], dtype=float), [A, B])),
                                  np.array(ka["expected"]))


def test_planted_dropped_weight_is_caught(known_answers):
    ka = known_answers["cp_to_tensor_weights"]
    A = np.arange(*ka["A_arange"][:2], dtype=float).reshape(ka["A_arange"][2:])
    B = np.arange(*ka["B_arange"][:2], dtype=float).reshape(ka["B_arange"][2:])
    assert not np.array_equal(T.cp_to_tensor((None, [A, B])), np.array(ka["expected"]))


def test_kinetic_fixture_decodes():
    """tests/golden/kinetic_slice.npz: the 32 x 12 x 10 x 30 slice of the
    reference's kinetic data, stored exactly as 3 x in int16; missing entries
    hold the zeros the reference's loader leaves."""
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kinetic_slice.npz"))
    x = d["x3"].astype(np.float64) / 3.0
    assert x.shape == (32, 12, 10, 30)
    assert d["missing"].shape == x.shape and d["missing"].any()
    assert np.all(x[d["missing"]] == 0.0)
    assert x.max() > 100.0
