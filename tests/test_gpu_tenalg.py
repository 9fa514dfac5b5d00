"""HIP tenalg kernels against the oracle (same seeded inputs), through the
C ABI.  Tolerances: FP64 paths 1e-12 relative; FP32 and BF16x3 MTTKRP 1e-5
relative (Frobenius) against the FP64 oracle; plain BF16 1e-2."""
import zlib

import numpy as np
import pytest
import torch

from oracle import tensorly_np as T
from conftest import arange_matrix

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def g(x, dev, dtype=torch.float64):
    return torch.tensor(np.asarray(x), dtype=dtype, device=dev)


def test_khatri_rao_known_answers(cuda, known_answers):
    import tensorly_amd as tla
    for key in ("kr_classic", "kr_eye"):
        ka = known_answers[key]
        mats = [g(m, cuda) for m in ka["inputs"]]
        np.testing.assert_array_equal(tla.khatri_rao(mats).cpu().numpy(), np.array(ka["expected"]))
    ka = known_answers["kr_arange4"]
    U = [g(arange_matrix(*spec), cuda) for spec in ka["inputs_arange"]]
    np.testing.assert_array_equal(tla.khatri_rao(U).cpu().numpy(), np.array(ka["expected"]))
    np.testing.assert_array_equal(tla.khatri_rao(U, skip_matrix=1).cpu().numpy(),
                                  tla.khatri_rao([U[0]] + U[2:]).cpu().numpy())
    assert tla.khatri_rao([U[0]]) is U[0]


def test_khatri_rao_weights_mask_float32(cuda):
    import tensorly_amd as tla
    rng = np.random.default_rng(0)
    mats = [rng.standard_normal((s, 5)) for s in (3, 4, 6)]
    w = rng.random(5)
    mask = (rng.random(72) > 0.3).astype(float)
    ref = T.khatri_rao(mats, weights=w, mask=mask)
    out = tla.khatri_rao([g(m, cuda, torch.float32) for m in mats], weights=g(w, cuda, torch.float32),
                         mask=g(mask, cuda, torch.float32))
    assert rel(out.cpu(), ref) < 1e-6


SHAPES = [(20, 24, 32), (16, 32, 64), (7, 9, 11), (12, 10, 8, 32), (5, 6, 7, 8), (40, 48, 128),
          (8, 16, 32, 64)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("rank", [1, 5, 16, 32, 64])
def test_mttkrp_fp64_valu(cuda, shape, rank):
    import tensorly_amd as tla
    rng = np.random.default_rng(zlib.crc32(repr((shape, rank)).encode()))
    X = rng.standard_normal(shape)
    w = rng.random(rank) + 0.5
    f = [rng.standard_normal((s, rank)) for s in shape]
    Xg, fg, wg = g(X, cuda), [g(x, cuda) for x in f], g(w, cuda)
    for mode in range(len(shape)):
        ref = T.unfolding_dot_khatri_rao(X, (w, f), mode)
        out = tla.unfolding_dot_khatri_rao(Xg, (wg, fg), mode)
        assert out.dtype == torch.float64
        assert rel(out.cpu(), ref) < 1e-12, (shape, rank, mode)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("rank", [3, 16, 32, 48, 64])
@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16x3", 1e-5), ("bf16", 1e-2), ("valu", 1e-5)])
def test_mttkrp_float32(cuda, shape, rank, prec, tol):
    import tensorly_amd as tla
    rng = np.random.default_rng(zlib.crc32(repr((shape, rank, prec)).encode()))
    X = rng.random(shape).astype(np.float32)
    w = (rng.random(rank) + 0.5).astype(np.float32)
    f = [rng.random((s, rank)).astype(np.float32) for s in shape]
    Xg, fg, wg = g(X, cuda, torch.float32), [g(x, cuda, torch.float32) for x in f], g(w, cuda, torch.float32)
    with tla.precision(prec):
        for mode in range(len(shape)):
            ref = T.unfolding_dot_khatri_rao(X.astype(np.float64), (w.astype(np.float64),
                                                                   [x.astype(np.float64) for x in f]), mode)
            out = tla.unfolding_dot_khatri_rao(Xg, (wg, fg), mode)
            assert out.dtype == torch.float32 and out.shape == ref.shape
            err = rel(out.cpu(), ref)
            assert err < tol, (shape, rank, prec, mode, err)


def test_mttkrp_signed_data_bf16x3(cuda):
    """Cancellation-heavy signed data: bf16x3 within 1e-5 (measured 4.2-4.6e-6
    on this case: the split operands carry 16 of FP32's 24 bits), plain BF16
    within 1e-2."""
    import tensorly_amd as tla
    rng = np.random.default_rng(5)
    shape, rank = (32, 48, 64), 32
    X = rng.standard_normal(shape).astype(np.float32)
    f = [rng.standard_normal((s, rank)).astype(np.float32) for s in shape]
    Xg, fg = g(X, cuda, torch.float32), [g(x, cuda, torch.float32) for x in f]
    for prec, tol in (("fp32", 1e-5), ("bf16x3", 1e-5), ("bf16", 1e-2)):
        with tla.precision(prec):
            for mode in range(3):
                ref = T.unfolding_dot_khatri_rao(X.astype(np.float64), (None, [x.astype(np.float64) for x in f]),
                                                 mode)
                err = rel(tla.unfolding_dot_khatri_rao(Xg, (None, fg), mode).cpu(), ref)
                assert err < tol, (prec, mode, err)


@pytest.mark.parametrize("shape", [(16, 32, 64), (8, 16, 24, 64)])
@pytest.mark.parametrize("rank", [16, 48, 64])
def test_mttkrp_bf16_copy(cuda, shape, rank):
    """A bfloat16 tensor is the BF16 copy of float32 data (MFMA path)."""
    import tensorly_amd as tla
    rng = np.random.default_rng(6 + rank)
    X = rng.random(shape).astype(np.float32)
    f = [rng.random((s, rank)).astype(np.float32) for s in shape]
    Xb = g(X, cuda, torch.float32).to(torch.bfloat16)
    fg = [g(x, cuda, torch.float32) for x in f]
    for mode in range(len(shape)):
        ref = T.unfolding_dot_khatri_rao(X.astype(np.float64), (None, [x.astype(np.float64) for x in f]), mode)
        err = rel(tla.unfolding_dot_khatri_rao(Xb, (None, fg), mode).cpu(), ref)
        assert err < 1e-2, (mode, err)


@pytest.mark.parametrize("shape", [(16, 32, 64), (8, 16, 24, 64), (40, 24, 96), (64, 64, 64, 64)])
@pytest.mark.parametrize("rank", [16, 33, 64])
@pytest.mark.parametrize("copy", [False, True])
def test_mttkrp_bf16_exact_for_rounded_tensor(cuda, shape, rank, copy):
    """BF16 multiplies the bf16-rounded tensor (rounded on the fly, or a
    bf16 copy) by the factors split in two bf16 terms: against the FP64
    oracle MTTKRP of the rounded tensor it is as accurate as bf16x3 (the
    split factors keep 16 of FP32's 24 bits), on signed data, for every
    kernel family these shapes reach (row-tile, last-mode, order 4)."""
    import tensorly_amd as tla
    rng = np.random.default_rng(zlib.crc32(repr((shape, rank, copy)).encode()))
    X = rng.standard_normal(shape).astype(np.float32)
    f = [rng.standard_normal((s, rank)).astype(np.float32) for s in shape]
    Xg = g(X, cuda, torch.float32)
    Xr = Xg.to(torch.bfloat16)
    X64 = Xr.double().cpu().numpy()
    fg = [g(x, cuda, torch.float32) for x in f]
    with tla.precision("bf16"):
        for mode in range(len(shape)):
            ref = T.unfolding_dot_khatri_rao(X64, (None, [x.astype(np.float64) for x in f]), mode)
            out = tla.unfolding_dot_khatri_rao(Xr if copy else Xg, (None, fg), mode)
            err = rel(out.cpu(), ref)
            assert err < 2e-5, (mode, err)


@pytest.mark.parametrize("fj", ["0", "1"])
@pytest.mark.parametrize("rank", [5, 16, 32, 64])
def test_last_mode_fj_staging(cuda, monkeypatch, fj, rank):
    """Last-mode kernel with the mode N-2 factor rows carried in registers
    (TLA_LAST_FJ=0) or staged in LDS (TLA_LAST_FJ=1), every rank tile count,
    against the FP64 oracle; both variants give the same bits."""
    import tensorly_amd as tla
    rng = np.random.default_rng(rank)
    shape = (24, 40, 200)
    X = rng.standard_normal(shape).astype(np.float32)
    f = [rng.standard_normal((s, rank)).astype(np.float32) for s in shape]
    Xg, fg = g(X, cuda, torch.float32), [g(x, cuda, torch.float32) for x in f]
    ref = T.unfolding_dot_khatri_rao(X.astype(np.float64), (None, [x.astype(np.float64) for x in f]), 2)
    for prec, tol in (("fp32", 1e-5), ("bf16x3", 1e-5), ("bf16", 1e-2)):
        with tla.precision(prec):
            monkeypatch.setenv("TLA_LAST_FJ", fj)
            out = tla.unfolding_dot_khatri_rao(Xg, (None, fg), 2)
            monkeypatch.setenv("TLA_LAST_FJ", "1" if fj == "0" else "0")
            other = tla.unfolding_dot_khatri_rao(Xg, (None, fg), 2)
        assert rel(out.cpu(), ref) < tol, (prec, rel(out.cpu(), ref))
        assert torch.equal(out, other), prec


@pytest.mark.parametrize("shape", [(32, 48, 64), (16, 24, 16, 64)])
def test_partial_mttkrp_bf16_copy(cuda, shape):
    """Dimension-tree partial contractions of a bfloat16 tensor (the BF16
    copy path, orders 3 and 4) against the FP64 einsum of the float32 data."""
    from tensorly_amd.tenalg import mttkrp_partial
    rng = np.random.default_rng(60 + len(shape))
    R = 16
    X = rng.standard_normal(shape).astype(np.float32)
    f = [rng.standard_normal((s, R)).astype(np.float32) for s in shape]
    Xb = g(X, cuda, torch.float32).to(torch.bfloat16)
    fg = [g(x, cuda, torch.float32) for x in f]
    X64, f64 = X.astype(np.float64), [x.astype(np.float64) for x in f]
    if len(shape) == 3:
        cases = {(0, 2): np.einsum("abc,cr->abr", X64, f64[2])}
    else:
        cases = {(0, 2): np.einsum("abcd,cr,dr->abr", X64, f64[2], f64[3]),
                 (2, 4): np.einsum("abcd,ar,br->cdr", X64, f64[0], f64[1])}
    for keep, ref in cases.items():
        got = mttkrp_partial(Xb, (None, fg), keep).cpu().numpy().reshape(ref.shape)
        assert np.all(np.isfinite(got))
        assert rel(got, ref) < 1e-2, (keep, rel(got, ref))


def test_mttkrp_deterministic(cuda):
    import tensorly_amd as tla
    rng = np.random.default_rng(7)
    X = g(rng.random((64, 64, 64)), cuda, torch.float32)
    f = [g(rng.random((64, 32)), cuda, torch.float32) for _ in range(3)]
    for mode in range(3):
        a = tla.unfolding_dot_khatri_rao(X, (None, f), mode)
        b = tla.unfolding_dot_khatri_rao(X, (None, f), mode)
        assert torch.equal(a, b)


@pytest.mark.parametrize("shape", [(16, 24, 32), (8, 16, 8, 32), (16, 24, 64), (8, 16, 16, 64)])
@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_partial_mttkrp_and_rankdiag(cuda, shape, prec):
    import tensorly_amd as tla
    from tensorly_amd.tenalg import mttkrp_partial, rankdiag_contract
    rng = np.random.default_rng(8)
    R = 16
    X = rng.random(shape).astype(np.float32)
    f = [rng.random((s, R)).astype(np.float32) for s in shape]
    X64, f64 = X.astype(np.float64), [x.astype(np.float64) for x in f]
    Xg, fg = g(X, cuda, torch.float32), [g(x, cuda, torch.float32) for x in f]
    with tla.precision(prec):
        T1 = mttkrp_partial(Xg, (None, fg), (0, 2)).reshape(shape[0], shape[1], R)
        # mode-0 / mode-1 MTTKRP from the shared partial contraction
        for mode, axis, F in ((0, 1, fg[1]), (1, 0, fg[0])):
            ref = T.unfolding_dot_khatri_rao(X64, (None, f64), mode)
            assert rel(rankdiag_contract(T1.contiguous(), F, axis).cpu(), ref) < 1e-5
        if len(shape) == 4:
            T2 = mttkrp_partial(Xg, (None, fg), (2, 4)).reshape(shape[2], shape[3], R)
            for mode, axis, F in ((2, 1, fg[3]), (3, 0, fg[2])):
                ref = T.unfolding_dot_khatri_rao(X64, (None, f64), mode)
                assert rel(rankdiag_contract(T2.contiguous(), F, axis).cpu(), ref) < 1e-5


@pytest.mark.parametrize("I,J,R", [(512, 48, 32), (300, 20, 8), (3, 5, 12), (64, 33, 7), (257, 64, 64)])
def test_rankdiag_contract_against_einsum(cuda, I, J, R):
    """Both contraction axes, vectorised (rank % 4 == 0, split reduction)
    and scalar forms, against the FP64 einsum."""
    from tensorly_amd.tenalg import rankdiag_contract
    rng = np.random.default_rng(I * 131 + J * 7 + R)
    Y = rng.standard_normal((I, J, R)).astype(np.float32)
    FJ = rng.standard_normal((J, R)).astype(np.float32)
    FI = rng.standard_normal((I, R)).astype(np.float32)
    Yg = g(Y, cuda, torch.float32)
    out1 = rankdiag_contract(Yg, g(FJ, cuda, torch.float32), 1).cpu().numpy()
    out0 = rankdiag_contract(Yg, g(FI, cuda, torch.float32), 0).cpu().numpy()
    ref1 = np.einsum("ijr,jr->ir", Y.astype(np.float64), FJ.astype(np.float64))
    ref0 = np.einsum("ijr,ir->jr", Y.astype(np.float64), FI.astype(np.float64))
    assert rel(out1, ref1) < 1e-5
    assert rel(out0, ref0) < 1e-5


def test_mode_dot_known_answers(cuda, known_answers):
    import tensorly_amd as tla
    ka = known_answers["mode_dot_matrix"]
    out = tla.mode_dot(g(ka["X"], cuda), g(ka["U"], cuda), ka["mode"])
    np.testing.assert_array_equal(out.cpu().numpy(), np.array(ka["expected"]))
    ka = known_answers["mode_dot_vector"]
    out = tla.mode_dot(g(ka["X"], cuda), g(ka["v"], cuda), ka["mode"])
    np.testing.assert_array_equal(out.cpu().numpy(), np.array(ka["expected"]))
    ka = known_answers["multi_mode_dot_vectors"]
    out = tla.multi_mode_dot(g(ka["X"], cuda), [g(u, cuda) for u in ka["U"]], ka["modes"])
    np.testing.assert_array_equal(out.reshape(-1).cpu().numpy(), np.array(ka["expected"]))


@pytest.mark.parametrize("shape,J", [((64, 96, 128), 32), ((32, 64, 256), 16), ((33, 20, 7), 5), ((16, 8, 32, 64), 48)])
@pytest.mark.parametrize("prec,tol", [("fp32", 1e-6), ("bf16x3", 1e-5), ("bf16", 1e-2)])
@pytest.mark.parametrize("transpose", [False, True])
def test_mode_dot_float32(cuda, shape, J, prec, tol, transpose):
    import tensorly_amd as tla
    rng = np.random.default_rng(zlib.crc32(repr((shape, J, prec, transpose)).encode()))
    X = rng.standard_normal(shape).astype(np.float32)
    with tla.precision(prec):
        for mode in range(len(shape)):
            U = rng.standard_normal((shape[mode], J) if transpose else (J, shape[mode])).astype(np.float32)
            ref = T.mode_dot(X.astype(np.float64), U.astype(np.float64), mode, transpose=transpose)
            out = tla.mode_dot(g(X, cuda, torch.float32), g(U, cuda, torch.float32), mode, transpose=transpose)
            assert tuple(out.shape) == ref.shape
            assert rel(out.cpu(), ref) < tol, (shape, J, prec, mode)


def test_multi_mode_dot_semantics(cuda):
    import tensorly_amd as tla
    rng = np.random.default_rng(9)
    dims = [2, 3, 4, 5]
    X = rng.standard_normal(dims)
    factors = [rng.random((dims[i] + 1, dims[i])) for i in range(4)]
    Xg, fg = g(X, cuda), [g(f, cuda) for f in factors]
    ref = T.multi_mode_dot(X, factors)
    assert rel(tla.multi_mode_dot(Xg, fg).cpu(), ref) < 1e-12
    ref = T.multi_mode_dot(X, factors, skip=1)
    assert rel(tla.multi_mode_dot(Xg, fg, skip=1).cpu(), ref) < 1e-12
    ref = T.multi_mode_dot(X, [f.T for f in factors], transpose=True)
    assert rel(tla.multi_mode_dot(Xg, [f.T.contiguous() for f in fg], transpose=True).cpu(), ref) < 1e-12
    vecs = [rng.random(dims[m]) for m in (3, 1, 0, 2)]
    ref = T.multi_mode_dot(X, vecs, modes=[3, 1, 0, 2])
    out = tla.multi_mode_dot(Xg, [g(v, cuda) for v in vecs], modes=[3, 1, 0, 2])
    assert out.dim() == 0
    np.testing.assert_allclose(float(out), float(ref), rtol=1e-12)


def test_cp_norm_and_to_tensor(cuda, known_answers):
    import tensorly_amd as tla
    ka = known_answers["cp_to_tensor_arange4"]
    U = [g(arange_matrix(*spec), cuda) for spec in ka["inputs_arange"]]
    np.testing.assert_array_equal(tla.cp_to_tensor((torch.ones(3, dtype=torch.float64, device=cuda), U)).cpu().numpy(),
                                  np.array(ka["expected"]))
    rng = np.random.default_rng(10)
    w = rng.random(6)
    f = [rng.standard_normal((s, 6)) for s in (8, 5, 7)]
    ref = T.cp_norm((w, f))
    out = tla.cp_norm((g(w, cuda), [g(x, cuda) for x in f]))
    np.testing.assert_allclose(float(out), ref, rtol=1e-12)


def test_unfold_fold_known_answers(cuda, known_answers):
    """The reference's unfolding table of arange(24) as 3 x 4 x 2
    (``tests/test_core.py``), on GPU tensors, and fold as its inverse."""
    import tensorly_amd as tla
    ka = known_answers["unfold_arange24_342"]
    X = torch.arange(24, dtype=torch.float64, device=cuda).reshape(ka["shape"])
    for mode, expected in enumerate(ka["unfoldings"]):
        U = tla.unfold(X, mode)
        assert U.device == X.device
        np.testing.assert_array_equal(U.cpu().numpy(), np.array(expected))
        back = tla.fold(U, mode, tuple(X.shape))
        assert torch.equal(back, X)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_mttkrp_order2_ignores_weights(cuda, dtype):
    """A matrix's MTTKRP is X @ B without the weights: the reference's
    khatri_rao of one matrix returns it unweighted (``_khatri_rao.py:67-69``,
    ``mttkrp.py:47-50``)."""
    import tensorly_amd as tla
    rng = np.random.default_rng(21)
    X = rng.standard_normal((40, 33))
    A, B = rng.standard_normal((40, 6)), rng.standard_normal((33, 6))
    w = rng.random(6) + 0.5
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    for mode in range(2):
        ref = T.unfolding_dot_khatri_rao(X, (w, [A, B]), mode)
        np.testing.assert_allclose(ref, (X @ B) if mode == 0 else (X.T @ A), rtol=1e-12)
        out = tla.unfolding_dot_khatri_rao(g(X, cuda, dtype), (g(w, cuda, dtype), [g(A, cuda, dtype),
                                                                                   g(B, cuda, dtype)]), mode)
        assert rel(out.cpu(), ref) < tol, mode


def test_operands_on_the_host_are_refused(cuda):
    """Every operand must be on the call's GPU: a host factor, weights or
    mask raises TypeError before any launch."""
    import tensorly_amd as tla
    rng = np.random.default_rng(22)
    X = g(rng.random((8, 9, 10)), cuda, torch.float32)
    f = [g(rng.random((s, 4)), cuda, torch.float32) for s in (8, 9, 10)]
    w = g(rng.random(4), cuda, torch.float32)
    host_f = [f[0], f[1].cpu(), f[2]]
    with pytest.raises(TypeError):
        tla.unfolding_dot_khatri_rao(X, (None, host_f), 0)
    with pytest.raises(TypeError):
        tla.unfolding_dot_khatri_rao(X, (w.cpu(), f), 0)
    with pytest.raises(TypeError):
        tla.khatri_rao(f, weights=w.cpu())
    with pytest.raises(TypeError):
        tla.khatri_rao(f, mask=torch.ones(8 * 9 * 10))
    with pytest.raises(TypeError):
        tla.khatri_rao([f[0], f[1].cpu()])
    with pytest.raises(TypeError):
        tla.mode_dot(X, f[0].T.contiguous().cpu(), 0)
    coords = torch.stack([torch.arange(5, device=cuda) % s for s in (8, 9, 10)])
    vals = torch.ones(5, device=cuda)
    with pytest.raises(TypeError):
        tla.tenalg.unfolding_dot_khatri_rao_coo(coords, vals, (8, 9, 10), (None, host_f), 0)
    with pytest.raises(TypeError):
        tla.tenalg.unfolding_dot_khatri_rao_coo(coords, vals, (8, 9, 10), (w.cpu(), f), 0)
