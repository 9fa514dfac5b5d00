"""The normal-equations update tla_cp_solve (the Gram-Hadamard solve of
tensorly/decomposition/_cp.py:411-427) through the C ABI against numpy FP64:
new factor rows M V^-1 (rounded to the factor dtype), their FP64 Gram and
<M, A>, for ranks on both sides of every padding size, partial last blocks,
weights, l2 regularisation and an indefinite V (partial-pivoting fallback)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _solve(cuda, grams, mode, M, weights=None, l2=0.0, dtype=torch.float32):
    from tensorly_amd import _lib
    L = _lib.lib()
    rows, R = M.shape
    gd = [torch.tensor(G, dtype=torch.float64, device=cuda) for G in grams]
    pp, keep = _lib.ptr_array([None if m == mode else g for m, g in enumerate(gd)])
    Mg = torch.tensor(M, dtype=dtype, device=cuda)
    wg = torch.tensor(weights, dtype=dtype, device=cuda) if weights is not None else None
    out = torch.empty((rows, R), dtype=dtype, device=cuda)
    gram = torch.empty((R, R), dtype=torch.float64, device=cuda)
    ip = torch.empty(1, dtype=torch.float64, device=cuda)
    status = torch.full((2,), -7, dtype=torch.int32, device=cuda)
    counter = torch.zeros(2, dtype=torch.int32, device=cuda)
    nb = L.tla_cp_solve_workspace_size(rows, R)
    ws = torch.empty(max(int(nb), 1), dtype=torch.uint8, device=cuda)
    code = _lib.F32 if dtype == torch.float32 else _lib.F64
    _lib.check(L.tla_cp_solve(code, pp, len(grams), mode, _lib.ptr(wg), float(l2), _lib.ptr(Mg), rows, R,
                              _lib.ptr(out), _lib.ptr(gram), _lib.ptr(ip), _lib.ptr(status), _lib.ptr(counter),
                              _lib.ptr(ws), nb, _lib.stream(cuda)), "cp_solve")
    torch.cuda.synchronize()
    assert int(counter[0]) == 0  # the finalising block leaves the counter at zero
    return (out.cpu().numpy(), gram.cpu().numpy(), float(ip[0]), int(status[0]))


def _reference(grams, mode, M, weights, l2, np_dtype):
    R = M.shape[1]
    V = np.ones((R, R))
    for m, G in enumerate(grams):
        if m != mode:
            V = V * G
    V = V + l2 * np.eye(R)
    if weights is not None:
        V = V * np.outer(weights, weights)
    A = np.linalg.solve(V, M.astype(np.float64).T).T
    Ar = A.astype(np_dtype).astype(np.float64)
    return A, Ar, V


@pytest.mark.parametrize("R", [1, 3, 5, 16, 17, 31, 32, 33, 48, 64])
@pytest.mark.parametrize("rows", [1, 64, 100, 512])
def test_cp_solve_against_numpy(cuda, R, rows):
    rng = np.random.default_rng(1000 * R + rows)
    facs = [rng.random((n, R)) + 0.05 for n in (40, 70, 55)]
    grams = [f.T @ f for f in facs]
    M = rng.standard_normal((rows, R)).astype(np.float32)
    A, Ar, V = _reference(grams, 1, M, None, 0.0, np.float32)
    out, gram, ip, status = _solve(cuda, grams, 1, M)
    assert status == 0
    cond = np.linalg.cond(V)
    # float32 output: rounding plus the FP64 solve's own cond(V) * eps error
    tol = 2e-7 + 1e-15 * cond
    assert np.max(np.abs(out - A)) <= tol * max(1.0, np.max(np.abs(A))), (R, rows, cond)
    np.testing.assert_allclose(gram, out.astype(np.float64).T @ out.astype(np.float64), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ip, np.sum(M.astype(np.float64) * out.astype(np.float64)), rtol=1e-11)


@pytest.mark.parametrize("R", [4, 24, 40, 64])
def test_cp_solve_float64_weights_l2(cuda, R):
    rng = np.random.default_rng(R)
    facs = [rng.standard_normal((n, R)) for n in (30, 90, 45, 20)]
    grams = [f.T @ f for f in facs]
    w = rng.random(R) + 0.5
    M = rng.standard_normal((130, R))
    A, _, V = _reference(grams, 2, M, w, 0.3, np.float64)
    out, gram, ip, status = _solve(cuda, grams, 2, M, weights=w, l2=0.3, dtype=torch.float64)
    assert status == 0
    np.testing.assert_allclose(out, A, rtol=0, atol=1e-13 * np.linalg.cond(V) * np.max(np.abs(A)))
    np.testing.assert_allclose(gram, out.T @ out, rtol=1e-12)
    np.testing.assert_allclose(ip, np.sum(M * out), rtol=1e-12)


@pytest.mark.parametrize("R", [6, 32, 50])
def test_cp_solve_indefinite_uses_pivoting(cuda, R):
    """A negative l2 term makes V indefinite: the SPD elimination meets a
    non-positive pivot and the partial-pivoting fallback (status 1) still
    returns M V^-1."""
    rng = np.random.default_rng(7 + R)
    facs = [rng.standard_normal((n, R)) for n in (20, 25, 30)]
    grams = [f.T @ f for f in facs]
    V0 = grams[1] * grams[2]
    shift = -float(np.median(np.linalg.eigvalsh(V0)))
    M = rng.standard_normal((70, R))
    A, _, V = _reference(grams, 0, M, None, shift, np.float64)
    assert np.min(np.linalg.eigvalsh(V)) < 0
    out, _, _, status = _solve(cuda, grams, 0, M, l2=shift, dtype=torch.float64)
    assert status == 1
    np.testing.assert_allclose(out, A, rtol=0, atol=1e-12 * np.linalg.cond(V) * np.max(np.abs(A)))
