"""The drop-ins on real data: a slice of the reference's kinetic
fluorescence tensor (tests/golden/kinetic_slice.npz, 32 x 12 x 10 x 30,
written by tests/golden/make_kinetic_fixture.py) against the oracle."""
import os

import numpy as np
import pytest
import torch

from oracle import tensorly_np as T

pytestmark = pytest.mark.gpu


def _kinetic():
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kinetic_slice.npz"))
    return d["x3"].astype(np.float64) / 3.0


@pytest.mark.parametrize("init", ["random", "svd"])
def test_parafac_on_kinetic_data(cuda, init):
    import tensorly_amd as tla
    X = _kinetic()
    cp, errs = tla.parafac(torch.tensor(X, device=cuda), 4, n_iter_max=30, init=init, random_state=0, tol=0,
                           return_errors=True)
    (w_r, f_r), errs_r = T.parafac(X, 4, n_iter_max=30, init=init, random_state=0, tol=0, return_errors=True)
    assert len(errs) == len(errs_r) == 30
    # random init: identical start; svd init: GPU Gram eigensolve vs LAPACK SVD
    np.testing.assert_allclose(errs, errs_r, rtol=1e-9 if init == "random" else 1e-6)
    rec = tla.cp_to_tensor(cp).cpu().numpy()
    ref = T.cp_to_tensor((w_r, f_r))
    assert np.linalg.norm(rec - ref) / np.linalg.norm(ref) < (1e-8 if init == "random" else 1e-5)


def test_tucker_on_kinetic_data(cuda):
    import tensorly_amd as tla
    X = _kinetic()
    rank = [6, 5, 4, 6]
    (core, factors), errs = tla.tucker(torch.tensor(X, device=cuda), rank, n_iter_max=10, init="svd", tol=0,
                                       return_errors=True)
    (core_r, fac_r), errs_r = T.tucker(X, rank, n_iter_max=10, init="svd", tol=0, return_errors=True)
    np.testing.assert_allclose(errs, errs_r, rtol=1e-8)
    rec = tla.tucker_to_tensor((core, factors)).cpu().numpy()
    ref = T.tucker_to_tensor((core_r, fac_r))
    assert np.linalg.norm(rec - ref) / np.linalg.norm(ref) < 1e-8


def test_non_negative_parafac_on_kinetic_data(cuda):
    import tensorly_amd as tla
    X = _kinetic()
    cp, errs = tla.non_negative_parafac(torch.tensor(X, device=cuda), 3, n_iter_max=25, init="random",
                                        random_state=1, tol=1e-300, return_errors=True)
    _, errs_r = T.non_negative_parafac(X, 3, n_iter_max=25, init="random", random_state=1, tol=1e-300,
                                       return_errors=True)
    np.testing.assert_allclose(errs, errs_r, rtol=1e-9)
    assert all(bool((f >= 0).all()) for f in cp.factors)
