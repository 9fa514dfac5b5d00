"""parafac / tucker drop-ins on the GPU against the oracle and the
reference trajectories (tests/golden/trajectories.npz)."""
import numpy as np
import pytest
import torch

from oracle import tensorly_np as T
from tensorly_amd import synthetic

pytestmark = pytest.mark.gpu


def g(x, dev, dtype=torch.float64):
    return torch.tensor(np.asarray(x), dtype=dtype, device=dev)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_parafac_fp64_matches_reference_trajectory(cuda, trajectories):
    """64^3 float64, rank 8, init='random' seed 0, 50 sweeps, tol=0."""
    import tensorly_amd as tla
    x, _ = synthetic.cp_64_rank8()
    cp, errs = tla.parafac(g(x, cuda), 8, n_iter_max=50, init="random", random_state=0, tol=0,
                           return_errors=True)
    np.testing.assert_allclose(errs, trajectories["c64_random_errors"], rtol=1e-7)
    rec = tla.cp_to_tensor(cp).cpu().numpy()
    ref = T.cp_to_tensor((trajectories["c64_random_weights"],
                          [trajectories[f"c64_random_factor{m}"] for m in range(3)]))
    assert rel(rec, ref) < 1e-6


@pytest.mark.parametrize("prec,dimtree", [("fp32", False), ("fp32", True), ("bf16x3", False), ("bf16x3", True)])
def test_parafac_float32_vs_oracle(cuda, prec, dimtree):
    """float32 tensor (64x48x96 rank-6 model + 1 % noise): per-sweep errors
    within 1e-5 of the FP64 oracle run on the same data and init."""
    import tensorly_amd as tla
    rng = np.random.default_rng(3)
    f_true = [rng.random((s, 6)) for s in (64, 48, 96)]
    x = synthetic.add_noise(synthetic.cp_model(np.ones(6), f_true, dtype=np.float32), 0.01, rng)
    (_, _), ref_errs = T.parafac(x.astype(np.float64), 6, n_iter_max=25, init="random", random_state=4, tol=0,
                                 return_errors=True)
    cp, errs = tla.parafac(g(x, cuda, torch.float32), 6, n_iter_max=25, init="random", random_state=4, tol=0,
                           return_errors=True, precision=prec, dimtree=dimtree)
    assert len(errs) == 25
    assert np.max(np.abs(np.array(errs) - np.array(ref_errs))) < 1e-5, (prec, dimtree)


# Seeds and tolerances where the FP64 oracle stops at sweep 9, 17 and 30, with
# the last two error decreases at least 1.5 % of tol on either side of it.
@pytest.mark.parametrize("tol,crit,seed", [(1e-3, "rec_error", 2), (1e-3, "abs_rec_error", 6),
                                           (3e-4, "rec_error", 1), (1e-7, "abs_rec_error", 6)])
def test_parafac_tol_chunked_matches_per_sweep_check(cuda, tol, crit, seed):
    """With only ``tol`` set, parafac checks the stopping rule once per chunk
    of sweeps and rolls back to the stopping sweep; a callback forces the
    per-sweep host check.  Both stop at the same sweep with the same bits,
    and, for tolerances the oracle reaches within 60 sweeps, at the sweep
    where the FP64 oracle stops."""
    import tensorly_amd as tla
    rng = np.random.default_rng(8)
    f_true = [rng.random((s, 5)) for s in (40, 36, 52)]
    x = synthetic.add_noise(synthetic.cp_model(np.ones(5), f_true, dtype=np.float32), 0.01, rng)
    xg = g(x, cuda, torch.float32)
    kw = dict(n_iter_max=60, init="random", random_state=seed, tol=tol, cvg_criterion=crit, return_errors=True)
    cp_a, errs_a = tla.parafac(xg, 5, **kw)
    cp_b, errs_b = tla.parafac(xg, 5, callback=lambda cp, e: False, **kw)
    assert errs_a == errs_b
    for fa, fb in zip(cp_a.factors, cp_b.factors):
        assert torch.equal(fa, fb)
    if tol >= 1e-4:
        (_, _), ref_errs = T.parafac(x.astype(np.float64), 5, **kw)
        assert len(errs_a) == len(ref_errs) < 60, (len(errs_a), len(ref_errs))


@pytest.mark.parametrize("shape", [(64, 48, 128), (16, 32, 16, 64)])
@pytest.mark.parametrize("prec,dimtree", [("fp32", False), ("bf16x3", True), ("fp32", True)])
@pytest.mark.parametrize("rank", [8, 40])
def test_parafac_persistent_kernels_vs_oracle(cuda, shape, prec, dimtree, rank):
    """Shapes whose MTTKRPs all run on the persistent streaming kernels, so
    every solve takes V^-1 from the side workgroup of the preceding MTTKRP
    launch or from the previous solve's finalising block: per-sweep err^2
    within 1e-6 of the FP64 oracle on the same data and init."""
    import tensorly_amd as tla
    rng = np.random.default_rng(rank + len(shape))
    f_true = [rng.random((s, rank)) for s in shape]
    x = synthetic.add_noise(synthetic.cp_model(np.ones(rank), f_true, dtype=np.float32), 0.01, rng)
    (_, _), ref_errs = T.parafac(x.astype(np.float64), rank, n_iter_max=12, init="random", random_state=5, tol=0,
                                 return_errors=True)
    _, errs = tla.parafac(g(x, cuda, torch.float32), rank, n_iter_max=12, init="random", random_state=5, tol=0,
                          return_errors=True, precision=prec, dimtree=dimtree)
    d = np.abs(np.array(errs) ** 2 - np.array(ref_errs) ** 2)
    assert np.max(d) < 1e-6, (prec, dimtree, d)


@pytest.mark.parametrize("prec,dimtree", [("fp32", False), ("bf16x3", True)])
def test_parafac_normalize_persistent_kernels(cuda, prec, dimtree):
    """normalize_factors=True (weights change every sweep) with the side
    workgroup forming V^-1 from the renormalised Grams and weights: errors
    and weights against the FP64 oracle."""
    import tensorly_amd as tla
    rng = np.random.default_rng(21)
    shape, rank = (64, 48, 128), 6
    f_true = [rng.random((s, rank)) for s in shape]
    x = synthetic.add_noise(synthetic.cp_model(np.ones(rank), f_true, dtype=np.float32), 0.01, rng)
    (w_ref, _), ref_errs = T.parafac(x.astype(np.float64), rank, n_iter_max=10, init="random", random_state=8,
                                     tol=0, normalize_factors=True, return_errors=True)
    cp, errs = tla.parafac(g(x, cuda, torch.float32), rank, n_iter_max=10, init="random", random_state=8, tol=0,
                           normalize_factors=True, return_errors=True, precision=prec, dimtree=dimtree)
    d = np.abs(np.array(errs) ** 2 - np.array(ref_errs) ** 2)
    assert np.max(d) < 1e-6, d
    np.testing.assert_allclose(cp.weights.cpu().numpy(), w_ref, rtol=1e-3)


@pytest.mark.parametrize("shape,rank", [((96, 80, 64), 32), ((96, 80, 64), 12), ((64, 72, 64), 20),
                                        ((24, 32, 64, 64), 16), ((64, 40, 72, 32), 28)])
def test_dimtree_sweep_matches_public_entry_points(cuda, shape, rank):
    """One sweep of the native dimension-tree driver against the same sweep
    composed from the public entry points, bit for bit, for every mode whose
    MTTKRP is a rank-diagonal contraction of a shared partial (all four
    modes of an order-4 tensor, modes 0 and 1 of an order-3 one).  The
    driver takes each V^-1 from the side workgroup of the partial MTTKRP
    launch or from the previous solve's finalising block; the public path
    is tla_mttkrp_partial, tla_rankdiag_contract and tla_cp_solve inverting
    V itself."""
    from tensorly_amd import _lib
    from tensorly_amd.cp_tensor import grams64
    from tensorly_amd.decomposition._cp import CpAlsEngine
    from tensorly_amd.tenalg import mttkrp_partial, rankdiag_contract
    rng = np.random.default_rng(sum(shape) * 7 + rank)
    x = torch.tensor(rng.random(shape), dtype=torch.float32, device=cuda)
    init = [torch.tensor(rng.random((s, rank)), dtype=torch.float32, device=cuda) for s in shape]
    eng = CpAlsEngine(x, torch.ones(rank, device=cuda), init, precision="bf16x3", dimtree=True, unit_weights=True)
    assert eng.dimtree
    eng.sweeps(1)
    N = len(shape)
    f = [a.clone() for a in init]
    G = grams64(f)
    L = _lib.lib()

    def solve(mode, m):
        rows = shape[mode]
        gp, _keep = _lib.ptr_array(G)
        g_out = torch.empty((rank, rank), dtype=torch.float64, device=cuda)
        status = torch.zeros(1, dtype=torch.int32, device=cuda)
        nb = L.tla_cp_solve_workspace_size(rows, rank)
        ws = _lib.workspace(nb, cuda)
        _lib.check(L.tla_cp_solve(_lib.F32, gp, N, mode, None, 0.0, _lib.ptr(m.contiguous()), rows, rank,
                                  _lib.ptr(f[mode]), _lib.ptr(g_out), None, _lib.ptr(status), None, _lib.ptr(ws),
                                  nb, _lib.stream(cuda)), "cp_solve")
        G[mode] = g_out
        assert int(status) == 0

    for kb in range(0, N - 1, 2):
        T = mttkrp_partial(x, (None, f), (kb, kb + 2), "bf16x3").reshape(shape[kb], shape[kb + 1], rank)
        solve(kb, rankdiag_contract(T, f[kb + 1], 1))
        solve(kb + 1, rankdiag_contract(T, f[kb], 0))
    modes = range(N) if N == 4 else (0, 1)
    for m in modes:
        d = (eng.factors[m] - f[m]).abs().max().item()
        assert torch.equal(eng.factors[m], f[m]), (shape, rank, m, d)


def test_parafac_order4_dimtree(cuda):
    import tensorly_amd as tla
    rng = np.random.default_rng(5)
    shape = (16, 24, 16, 32)
    f_true = [rng.standard_normal((s, 5)) for s in shape]
    x = synthetic.add_noise(synthetic.cp_model(np.ones(5), f_true, dtype=np.float32), 0.01, rng)
    (_, _), ref_errs = T.parafac(x.astype(np.float64), 5, n_iter_max=20, init="random", random_state=2, tol=0,
                                 return_errors=True)
    for dimtree in (False, True):
        _, errs = tla.parafac(g(x, cuda, torch.float32), 5, n_iter_max=20, init="random", random_state=2, tol=0,
                              return_errors=True, dimtree=dimtree)
        assert np.max(np.abs(np.array(errs) - np.array(ref_errs))) < 1e-5


@pytest.mark.parametrize("shape", [(48, 40, 64), (16, 24, 16, 32)])
@pytest.mark.parametrize("dimtree", [False, True])
def test_parafac_bf16_copy_tracks_fp64(cuda, shape, dimtree):
    """ALS whose MTTKRPs read a bfloat16 copy of the tensor (BF16 MFMA, FP32
    accumulation; ||X||^2 from the float32 data).  The factors it returns
    reconstruct X to within 1e-5 (absolute, relative error) of the FP64
    oracle's final error.  The reported error is the cheap form
    sqrt(||X||^2 + ||M||^2 - 2 <X, M>) with <X, M> from the BF16 MTTKRP,
    whose relative error (~1e-4 here) is amplified by the cancellation, so it
    is checked on err^2 with a 1e-3 tolerance."""
    import tensorly_amd as tla
    rng = np.random.default_rng(11 + len(shape))
    f_true = [rng.standard_normal((s, 5)) for s in shape]
    x = synthetic.add_noise(synthetic.cp_model(np.ones(5), f_true, dtype=np.float32), 0.01, rng)
    (_, _), ref_errs = T.parafac(x.astype(np.float64), 5, n_iter_max=30, init="random", random_state=3, tol=0,
                                 return_errors=True)
    xg = g(x, cuda, torch.float32)
    cp, errs = tla.parafac(xg, 5, n_iter_max=30, init="random", random_state=3, tol=0, return_errors=True,
                           precision="bf16", tensor_bf16=xg.to(torch.bfloat16), dimtree=dimtree)
    errs = np.array(errs)
    assert np.all(np.isfinite(errs))
    x64 = x.astype(np.float64)
    rec = T.cp_to_tensor((cp.weights.double().cpu().numpy(), [f.double().cpu().numpy() for f in cp.factors]))
    true_err = np.linalg.norm(x64 - rec) / np.linalg.norm(x64)
    assert abs(true_err - ref_errs[-1]) < 1e-5, (true_err, ref_errs[-1])
    assert abs(errs[-1] ** 2 - ref_errs[-1] ** 2) < 1e-3, (errs[-1], ref_errs[-1])


def test_parafac_bf16_copy_rejects_unsupported_shape(cuda):
    """A bfloat16 copy has no generic MTTKRP: shapes outside the MFMA plans
    (last mode not a multiple of 32) are rejected when the driver is made."""
    import tensorly_amd as tla
    xg = torch.rand((16, 24, 56), device=cuda)
    with pytest.raises(ValueError, match="bf16 tensor copy"):
        tla.parafac(xg, 4, n_iter_max=2, init="random", random_state=0, precision="bf16",
                    tensor_bf16=xg.to(torch.bfloat16))


def test_parafac_linesearch_matches_reference(cuda, trajectories):
    import tensorly_amd as tla
    x, _ = synthetic.cp_64_rank8()
    _, errs = tla.parafac(g(x, cuda), 8, n_iter_max=30, init="random", random_state=0, tol=0, linesearch=True,
                          return_errors=True)
    np.testing.assert_allclose(errs, trajectories["c64_ls_errors"], rtol=1e-7)


def test_parafac_svd_init_normalize_tol(cuda, trajectories):
    import tensorly_amd as tla
    x4 = trajectories["x4"]
    w0, f0 = tla.decomposition.initialize_cp(g(x4, cuda), 5, init="svd")
    for m in range(4):
        np.testing.assert_allclose(f0[m].cpu().numpy(), trajectories[f"x4_svdinit_factor{m}"], rtol=1e-6,
                                   atol=1e-8)
    cp, errs = tla.parafac(g(x4, cuda), 5, n_iter_max=40, init="svd", tol=1e-9, normalize_factors=True,
                           return_errors=True)
    ref = trajectories["x4_svd_norm_errors"]
    assert len(errs) == len(ref)
    np.testing.assert_allclose(errs, ref, rtol=1e-6)
    np.testing.assert_allclose(cp.weights.cpu().numpy(), trajectories["x4_svd_norm_weights"], rtol=1e-5)


def test_parafac_fixed_modes_l2_callback(cuda):
    import tensorly_amd as tla
    rng = np.random.default_rng(6)
    x = rng.random((12, 16, 20))
    seen = []

    def cb(cp, err):
        seen.append(err)
        return len(seen) == 4

    ref_cp, ref_errs = T.parafac(x, 3, n_iter_max=10, init="random", random_state=1, tol=0, l2_reg=0.1,
                                 fixed_modes=[0], return_errors=True)
    cp, errs = tla.parafac(g(x, cuda), 3, n_iter_max=10, init="random", random_state=1, tol=0, l2_reg=0.1,
                           fixed_modes=[0], return_errors=True)
    np.testing.assert_allclose(errs, ref_errs, rtol=1e-9)
    init0 = np.random.RandomState(1).random_sample((12, 3))
    np.testing.assert_array_equal(cp.factors[0].cpu().numpy(), init0)
    _ = tla.parafac(g(x, cuda), 3, n_iter_max=10, init="random", random_state=1, tol=0, callback=cb)
    assert len(seen) == 4


@pytest.mark.parametrize("n,k", [(40, 5), (64, 64), (200, 32), (256, 40), (300, 41), (256, 56), (100, 57),
                                 (128, 64), (128, 128), (200, 100), (300, 120)])
def test_gram_eigensolver(cuda, n, k):
    from tensorly_amd.svd import gram_eigh
    rng = np.random.default_rng(n)
    # spectrum with a gap after k (like a Tucker model plus noise)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    lam = np.concatenate([np.sort(rng.random(k) + 1.0)[::-1] * 10, rng.random(n - k) * 0.5])
    G = (Q * lam) @ Q.T
    evals, evecs = gram_eigh(g(G, cuda), k)
    ref_vals, ref_vecs = np.linalg.eigh(G)
    ref_vals, ref_vecs = ref_vals[::-1][:k], ref_vecs[:, ::-1][:, :k]
    np.testing.assert_allclose(evals.cpu().numpy(), ref_vals, rtol=1e-10)
    overlap = np.abs(np.sum(evecs.cpu().numpy() * ref_vecs, axis=0))
    np.testing.assert_allclose(overlap, np.ones(k), atol=1e-8)


def test_gram_eigensolver_batch_and_warm_start(cuda):
    """tla_sym_eig_batch (one workgroup per problem, problems of every
    solver kind in one call) returns what single calls return, bit for bit;
    an FP64 warm start (the converged vectors, as the HOOI loop passes them)
    reaches the same eigenpairs."""
    from tensorly_amd.svd import gram_eigh, gram_eigh_batch
    cases = [(40, 5), (100, 20), (256, 32), (300, 45), (256, 32), (200, 100)]
    Gs, refs = [], []
    for i, (n, k) in enumerate(cases):
        rng = np.random.default_rng(100 + i)
        Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
        lam = np.concatenate([np.sort(rng.random(k) + 1.0)[::-1] * 10, rng.random(n - k) * 0.5])
        G = (Q * lam) @ Q.T
        Gs.append(g(G, cuda))
        rv, rvec = np.linalg.eigh(G)
        refs.append((rv[::-1][:k], rvec[:, ::-1][:, :k]))
    single = [gram_eigh(G, k) for G, (_, k) in zip(Gs, cases)]
    batch = gram_eigh_batch(Gs, [k for _, k in cases])
    for (e1, v1), (e2, v2) in zip(single, batch):
        assert torch.equal(e1, e2) and torch.equal(v1, v2)
    for (G, (n, k)), (ev, vec), (rv, rvec) in zip(zip(Gs, cases), single, refs):
        e3, v3 = gram_eigh(G, k, warm=vec)
        np.testing.assert_allclose(e3.cpu().numpy(), rv, rtol=1e-10)
        overlap = np.abs(np.sum(v3.cpu().numpy() * rvec, axis=0))
        np.testing.assert_allclose(overlap, np.ones(k), atol=1e-8)
        orth = v3.cpu().numpy().T @ v3.cpu().numpy()
        np.testing.assert_allclose(orth, np.eye(k), atol=1e-11)


def _svd_flip_np(U):
    """The reference's u-based sign rule (tenalg/svd.py:13-42) in numpy."""
    idx = np.argmax(np.abs(U), axis=0)
    return U * np.sign(U[idx, np.arange(U.shape[1])])


@pytest.mark.parametrize("n,k", [(40, 5), (100, 20), (256, 32), (300, 45), (200, 100)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gram_left_singular_fused(cuda, n, k, dtype):
    """tla_gram_left_singular: the leading left singular vectors of a matrix
    A from its Gram, with svd_flip's signs, against numpy's SVD of A; the
    same vectors (bit for bit, before the cast) as tla_sym_eig followed by
    the sign rule on the host; and the returned block warm-starts solves on
    perturbed matrices (1e-4, then 1e-8) that land on their own SVDs."""
    from tensorly_amd.svd import gram_eigh, gram_left_singular
    rng = np.random.default_rng(7 * n + k)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    sv = np.concatenate([np.sort(rng.random(k) + 1.0)[::-1] * 3, rng.random(n - k) * 0.3])
    A = (Q * sv) @ np.linalg.qr(rng.standard_normal((n + 20, n)))[0].T  # n x (n + 20)
    G = A @ A.T
    U, S, blk = gram_left_singular(g(G, cuda), k, dtype, want_block=True)
    assert U.dtype == dtype and S.dtype == dtype
    Ur, Sr, _ = np.linalg.svd(A, full_matrices=False)
    Ur = _svd_flip_np(Ur[:, :k])
    tol = 1e-6 if dtype == torch.float32 else 1e-9
    np.testing.assert_allclose(S.cpu().numpy(), Sr[:k], rtol=tol)
    np.testing.assert_allclose(U.cpu().numpy(), Ur, atol=tol)
    ev, vec = gram_eigh(g(G, cuda), k)
    flipped = _svd_flip_np(vec.cpu().numpy())
    np.testing.assert_array_equal(U.cpu().numpy(), flipped.astype(U.cpu().numpy().dtype))
    if n <= 128:
        assert blk is None
        return
    assert blk.shape[0] == n and blk.shape[1] >= k + 8
    np.testing.assert_allclose(blk.cpu().numpy().T @ blk.cpu().numpy(), np.eye(blk.shape[1]), atol=1e-12)
    A2 = A + 1e-4 * rng.standard_normal(A.shape)
    G2 = A2 @ A2.T
    U2, S2, blk2 = gram_left_singular(g(G2, cuda), k, dtype, warm=blk, want_block=True)
    assert blk2.data_ptr() == blk.data_ptr()  # the block is updated in place
    Ur2, Sr2, _ = np.linalg.svd(A2, full_matrices=False)
    np.testing.assert_allclose(S2.cpu().numpy(), Sr2[:k], rtol=tol)
    np.testing.assert_allclose(U2.cpu().numpy(), _svd_flip_np(Ur2[:, :k]), atol=tol)
    # a much smaller move (the leading Ritz block is then diagonal to ~1e-8,
    # the case its perturbation-theory eigenvectors take)
    A3 = A2 + 1e-8 * rng.standard_normal(A.shape)
    U3, S3, _ = gram_left_singular(g(A3 @ A3.T, cuda), k, dtype, warm=blk2, want_block=True)
    Ur3, Sr3, _ = np.linalg.svd(A3, full_matrices=False)
    np.testing.assert_allclose(S3.cpu().numpy(), Sr3[:k], rtol=tol)
    np.testing.assert_allclose(U3.cpu().numpy(), _svd_flip_np(Ur3[:, :k]), atol=tol)


def test_tucker_core_equals_full_projection(cuda):
    """HOOI forms each iteration's core from the last mode's projected
    tensor; with ascending modes that is the same chain of products as the
    reference's multi_mode_dot over all modes, bit for bit (also for a
    subset of the modes).  With unsorted modes the reference's skip index
    (a position in the mode-sorted list) pairs the wrong factors and the
    core product raises, as it does here."""
    import tensorly_amd as tla
    from tensorly_amd.tenalg import multi_mode_dot
    rng = np.random.default_rng(3)
    x = g(rng.standard_normal((30, 140, 24)).astype(np.float32), cuda)
    (core, factors), _ = tla.partial_tucker(x, [5, 12, 6], n_iter_max=3, init="svd", tol=0)
    assert torch.equal(core, multi_mode_dot(x, factors, modes=[0, 1, 2], transpose=True))
    (core, factors), _ = tla.partial_tucker(x, [5, 6], modes=[0, 2], n_iter_max=3, init="svd", tol=0)
    assert torch.equal(core, multi_mode_dot(x, factors, modes=[0, 2], transpose=True))
    with pytest.raises(ValueError, match="not aligned"):
        tla.partial_tucker(x, [12, 5], modes=[1, 0], n_iter_max=1, init="svd", tol=0)


def test_gram_eigensolver_rejects_large_k(cuda):
    from tensorly_amd.svd import gram_eigh
    G = torch.eye(200, dtype=torch.float64, device=cuda)
    with pytest.raises(NotImplementedError, match="not supported"):
        gram_eigh(G, 121)


def test_svd_init_rank_above_56(cuda):
    """init='svd' at rank 64 on 128-wide modes (the 64-column block would
    hold at most 56 converged pairs) against the oracle's SVD init; the
    24-wide mode is padded with the same RandomState draws."""
    import tensorly_amd as tla
    rng = np.random.default_rng(8)
    x = rng.standard_normal((128, 96, 24))
    _, f = tla.decomposition.initialize_cp(g(x, cuda), 64, init="svd", random_state=0)
    _, f_ref = T.initialize_cp(x, 64, init="svd", random_state=0)
    for m in range(3):
        np.testing.assert_allclose(f[m].cpu().numpy(), f_ref[m], atol=1e-8)


def test_svd_interface_sign_rule(cuda):
    import tensorly_amd.svd as S
    rng = np.random.default_rng(7)
    A = rng.standard_normal((30, 50))
    U, s, V = S.svd_interface(g(A, cuda), n_eigenvecs=6)
    Ur, sr, Vr = T.svd_interface(A, n_eigenvecs=6)
    np.testing.assert_allclose(s.cpu().numpy(), sr, rtol=1e-10)
    np.testing.assert_allclose(U.cpu().numpy(), Ur, atol=1e-9)
    np.testing.assert_allclose(V.cpu().numpy(), Vr, atol=1e-9)


def test_tucker_matches_reference_trajectory(cuda, trajectories):
    import tensorly_amd as tla
    xt = trajectories["xt"]
    tt, errs = tla.tucker(g(xt, cuda), rank=[4, 5, 3], n_iter_max=10, tol=0, return_errors=True)
    np.testing.assert_allclose(errs, trajectories["xt_errors"], rtol=1e-9)
    np.testing.assert_allclose(tt.core.cpu().numpy(), trajectories["xt_core"], rtol=1e-7, atol=1e-9)
    for m in range(3):
        np.testing.assert_allclose(tt.factors[m].cpu().numpy(), trajectories[f"xt_factor{m}"], atol=1e-8)


@pytest.mark.parametrize("prec,tol_e2,tol_core", [("fp32", 5e-8, 1e-5), ("bf16x3", 2e-6, 1e-4)])
def test_tucker_float32_vs_oracle(cuda, prec, tol_e2, tol_core):
    """The reference error sqrt(|‖X‖² − ‖core‖²|)/‖X‖ cancels: at 1 % noise
    err² ≈ 1e-4, so the check is on err² (i.e. on ‖core‖²/‖X‖²), with the
    contraction's own relative accuracy as tolerance."""
    import tensorly_amd as tla
    x, _ = synthetic.tucker_256_core32(dtype=np.float32, size=96, rank=12)
    (core, fac), ref_errs = T.tucker(x.astype(np.float64), rank=[12, 12, 12], n_iter_max=5, tol=0,
                                     return_errors=True)
    tt, errs = tla.tucker(g(x, cuda, torch.float32), rank=[12, 12, 12], n_iter_max=5, tol=0, return_errors=True,
                          precision=prec)
    d = np.abs(np.array(errs) ** 2 - np.array(ref_errs) ** 2)
    assert np.max(d) < tol_e2, d
    for m in range(3):
        # principal angles between subspaces
        s = np.linalg.svd(fac[m].T @ tt.factors[m].cpu().numpy().astype(np.float64), compute_uv=False)
        assert np.min(s) > 1 - 1e-6
    # the core in the reference's basis (the sign rule makes the bases agree)
    assert rel(tt.core.cpu().numpy(), core) < tol_core


def _duplicated_init(rng, shape, rank):
    """Random factors whose columns 0 and 1 are equal in every mode: each
    Gram, their Hadamard product and so V then have two equal rows."""
    f = [rng.random((s, rank)) for s in shape]
    for a in f:
        a[:, 1] = a[:, 0]
    return np.ones(rank), f


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("tol", [0, 1e-8])
def test_parafac_singular_system_raises(cuda, dtype, tol):
    """Duplicated factor columns make the normal equations exactly singular:
    the reference's np.linalg.solve raises LinAlgError (``_cp.py:425-427``),
    and so does the drop-in, on the no-sync path (checked at return) and on
    the host loop (checked per sweep)."""
    import tensorly_amd as tla
    rng = np.random.default_rng(31)
    shape, rank = (20, 24, 28), 4
    x = rng.random(shape)
    w0, f0 = _duplicated_init(rng, shape, rank)
    with pytest.raises(np.linalg.LinAlgError):
        T.parafac(x, rank, n_iter_max=3, init=(w0, [a.copy() for a in f0]), tol=tol)
    with pytest.raises(np.linalg.LinAlgError):
        tla.parafac(g(x, cuda, dtype), rank, n_iter_max=3, init=(g(w0, cuda, dtype), [g(a, cuda, dtype) for a in f0]),
                    tol=tol)


def test_parafac_coo_singular_system_raises(cuda):
    """The sparse drop-in raises as the dense one on duplicated columns; a
    well-posed problem from a random init does not."""
    import tensorly_amd as tla
    rng = np.random.default_rng(32)
    shape, rank = (12, 14, 16), 3
    dense = rng.random(shape) * (rng.random(shape) < 0.3)
    coords = torch.tensor(np.stack(np.nonzero(dense)), device=cuda)
    vals = torch.tensor(dense[np.nonzero(dense)], device=cuda)
    w0, f0 = _duplicated_init(rng, shape, rank)
    with pytest.raises(np.linalg.LinAlgError):
        T.parafac(dense, rank, n_iter_max=2, init=(w0, [a.copy() for a in f0]), tol=0)
    with pytest.raises(np.linalg.LinAlgError):
        tla.parafac_coo(coords, vals, shape, rank, n_iter_max=2, init=(g(w0, cuda), [g(a, cuda) for a in f0]), tol=0)
    tla.parafac_coo(coords, vals, shape, rank, n_iter_max=4, init="random", random_state=0, tol=0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_parafac_matrix_normalize_factors(cuda, dtype):
    """Order 2 with normalize_factors=True: the weights scale V but not the
    MTTKRP (khatri_rao of one matrix is unweighted, ``_khatri_rao.py:67-69``).
    The first sweep (unit weights) and its normalisation match the oracle.
    From then on the reference's iteration is unstable: the second sweep's
    normal equations have condition numbers of 1e9 and 1e16, so its values are
    set by rounding (numpy's LU and the GPU's elimination differ by tens of
    percent) and both errors jump from 0.79 to above 1e4."""
    import tensorly_amd as tla
    rng = np.random.default_rng(33)
    M = rng.random((30, 4)) @ rng.random((4, 45)) + 0.01 * rng.standard_normal((30, 45))
    (w_ref, f_ref), e_ref = T.parafac(M, 4, n_iter_max=1, init="random", random_state=2, tol=0,
                                      normalize_factors=True, return_errors=True)
    cp, errs = tla.parafac(g(M, cuda, dtype), 4, n_iter_max=1, init="random", random_state=2, tol=0,
                           normalize_factors=True, return_errors=True)
    rtol = 1e-9 if dtype == torch.float64 else 1e-3
    np.testing.assert_allclose(errs, e_ref, rtol=rtol)
    np.testing.assert_allclose(cp.weights.cpu().numpy(), w_ref, rtol=rtol)
    for a, b in zip(cp.factors, f_ref):
        np.testing.assert_allclose(a.cpu().numpy(), b, rtol=10 * rtol, atol=10 * rtol)
        np.testing.assert_allclose(np.linalg.norm(a.cpu().numpy(), axis=0), 1.0, rtol=rtol)
    # the second sweep's first V = (A_1^T A_1) * w w^T is already badly
    # conditioned (its second one near 1e16)
    V = (f_ref[1].T @ f_ref[1]) * np.outer(w_ref, w_ref)
    assert np.linalg.cond(V) > 1e8
    _, e2_ref = T.parafac(M, 4, n_iter_max=2, init="random", random_state=2, tol=0, normalize_factors=True,
                          return_errors=True)
    _, e2 = tla.parafac(g(M, cuda, dtype), 4, n_iter_max=2, init="random", random_state=2, tol=0,
                        normalize_factors=True, return_errors=True)
    assert e2_ref[1] > 1e4 and e2[1] > 1e4


@pytest.mark.parametrize("shape", [(10, 3), (64, 64), (300, 17), (1000, 48), (5, 5)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_qr_matches_numpy_signs(cuda, shape, dtype):
    """tla_qr against numpy's qr (LAPACK geqrf + orgqr): the same Q and R,
    signs included, and Q R = A."""
    from tensorly_amd.svd import qr
    rng = np.random.default_rng(shape[0] + shape[1])
    A = rng.standard_normal(shape)
    Q, R = qr(g(A, cuda, dtype))
    Qr, Rr = np.linalg.qr(A)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    np.testing.assert_allclose(Q.cpu().numpy(), Qr, atol=tol * 10)
    np.testing.assert_allclose(R.cpu().numpy(), Rr, atol=tol * 10 * np.abs(Rr).max())
    np.testing.assert_allclose((Q.double() @ R.double()).cpu().numpy(), A, atol=tol * 10 * np.abs(A).max())


@pytest.mark.parametrize("k", [True, 4])
def test_parafac_orthogonalise_matches_reference(cuda, k):
    """orthogonalise=True QRs the factors before sweeps 0 and 1 (True is an
    int), an int k before sweeps 0..k: the drop-in follows the reference's
    recorded run (tests/golden/more_trajectories.npz)."""
    import os
    import tensorly_amd as tla
    more = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "more_trajectories.npz"))
    cp, errs = tla.parafac(g(more["ox"], cuda), 3, n_iter_max=8, init="random", random_state=5, tol=0,
                           orthogonalise=k, return_errors=True)
    np.testing.assert_allclose(errs, more[f"orth_{int(k)}_errors"], rtol=1e-9)
    for m in range(3):
        np.testing.assert_allclose(cp.factors[m].cpu().numpy(), more[f"orth_{int(k)}_factor{m}"], rtol=1e-7,
                                   atol=1e-10)


@pytest.mark.parametrize("init", ["random", "svd"])
def test_parafac_mask_matches_reference(cuda, init):
    """mask: the missing entries are re-imputed from the model after every
    sweep and the error is that of the imputed tensor; init='svd' imputes the
    unfoldings first (svd_mask_repeats=3).  Against the reference's recorded
    runs (tests/golden/more_trajectories.npz)."""
    import os
    import tensorly_amd as tla
    more = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "more_trajectories.npz"))
    kw = dict(init="random", random_state=6) if init == "random" else dict(init="svd", svd_mask_repeats=3)
    X = g(more["mx"], cuda)
    cp, errs = tla.parafac(X, 3, n_iter_max=10, tol=0, mask=g(more["mmask"], cuda), return_errors=True, **kw)
    np.testing.assert_allclose(errs, more[f"mask_{init}_errors"], rtol=1e-8)
    for m in range(3):
        np.testing.assert_allclose(cp.factors[m].cpu().numpy(), more[f"mask_{init}_factor{m}"], rtol=1e-6,
                                   atol=1e-9)
    # the caller's tensor is not modified
    np.testing.assert_array_equal(X.cpu().numpy(), more["mx"])


@pytest.mark.parametrize("linesearch", [False, True])
def test_parafac_mask_float32_vs_oracle(cuda, linesearch):
    """float32, rank 5, 30 % missing, with and without the line search,
    against the FP64 oracle: err^2 within 1e-6 at every recorded sweep."""
    import tensorly_amd as tla
    rng = np.random.default_rng(60)
    f = [rng.random((s, 5)) for s in (40, 36, 64)]
    x = synthetic.add_noise(synthetic.cp_model(np.ones(5), f, dtype=np.float32), 0.01, rng)
    mask = (rng.random(x.shape) > 0.3).astype(np.float32)
    (_, _), ref = T.parafac(x.astype(np.float64), 5, n_iter_max=20, init="random", random_state=3, tol=0,
                            mask=mask.astype(np.float64), linesearch=linesearch, return_errors=True)
    _, errs = tla.parafac(g(x, cuda, torch.float32), 5, n_iter_max=20, init="random", random_state=3, tol=0,
                          mask=g(mask, cuda, torch.float32), linesearch=linesearch, return_errors=True)
    assert len(errs) == len(ref)
    d = np.abs(np.array(errs) ** 2 - np.array(ref) ** 2)
    assert np.max(d) < 1e-6, d


def test_error_calc_with_mask(cuda):
    import tensorly_amd as tla
    from tensorly_amd.decomposition import error_calc
    rng = np.random.default_rng(61)
    x = rng.random((7, 8, 9))
    mask = (rng.random(x.shape) > 0.4).astype(np.float64)
    w = rng.random(3)
    f = [rng.random((s, 3)) for s in x.shape]
    e_ref, x_ref, n_ref = T.error_calc(x, np.linalg.norm(x), w, f, mask=mask)
    e, xt, n = error_calc(g(x, cuda), float(np.linalg.norm(x)), g(w, cuda), [g(a, cuda) for a in f],
                          mask=g(mask, cuda))
    np.testing.assert_allclose(float(e), e_ref, rtol=1e-12)
    np.testing.assert_allclose(float(n), n_ref, rtol=1e-12)
    np.testing.assert_allclose(xt.cpu().numpy(), x_ref, rtol=1e-13)
