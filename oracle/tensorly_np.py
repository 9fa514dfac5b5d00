"""Plain-numpy FP64 restatement of the reference TensorLy 0.9.0 CP-ALS /
Tucker-HOOI path.

TEST INFRASTRUCTURE ONLY -- the checker for the HIP path, never the thing
measured or shipped.  Product code must not import it.

Every function cites the reference file:line whose numerics it follows
(paths relative to the reference repository root, ``tensorly/...``).  The op
sequence mirrors the reference so that the CPU baseline in ``bench.py`` times
the same work (``unfold`` copy, Khatri-Rao materialisation, ``np.dot``,
Hadamard of Grams, ``np.linalg.solve``, ``error_calc``).

Pinning (how each function is checked, see ``tests/test_oracle_*.py``):

* ``unfold``/``fold``/``khatri_rao``/``mode_dot``/``multi_mode_dot``/
  ``cp_to_tensor``: the reference's own known-answer tables
  (``tests/golden/known_answers.json``) plus identities.
* ``unfolding_dot_khatri_rao`` (+ memory and einsum forms): brute-force
  quadruple loop on small inputs, and the three forms against each other.
* ``cp_norm``: brute-force ``||cp_to_tensor||``.  ``cp_normalize``: unit
  columns and unchanged reconstruction.
* ``error_calc``: brute-force ``||X - cp_to_tensor||``.
* ``random_cp`` / ``initialize_cp(init='random')``: the RandomState draw order,
  checked against the reference run recorded in the trajectory fixtures.
* ``parafac``, ``initialize_cp(init='svd')``, ``partial_tucker``/``tucker``:
  per-iteration trajectories recorded from the reference implementation
  (``tests/golden/trajectories.npz``, written by
  ``tests/golden/make_trajectories.py``), plus monotone-decrease and
  least-squares optimality (gradient = 0 after each mode update).
* ``svd_flip``/``truncated_svd``/``symeig_svd``: LAPACK identities
  (``U S V = A``, orthonormality) and the sign rule checked on hand-built cases.
* ``non_negative_parafac``: the Lee-Seung monotone decrease of the error,
  exact factorisations (with non-unit weights) as fixed points of the update,
  and a perturbed factor moving towards the solution
  (``tests/test_oracle_nn_cp.py``).
* ``hals_nnls`` / ``non_negative_parafac_hals``: the NNLS KKT conditions at
  the ``exact`` solution, the monotone decrease of the error, and exact
  factorisations as fixed points.
* ``unfolding_dot_khatri_rao_coo``: the dense MTTKRP of the densified tensor
  (repeated coordinates summed).
* ``cp_lstsq_grad``: central finite differences of the loss for every factor
  entry of a small model, and the loss against ``0.5 ||X - cp_to_tensor||^2``.
* ``make_svd_non_negative``: a non-negative rank-1 matrix reproduced exactly,
  non-negative outputs, the leading triplet kept, the NNDSVDa fill value, and
  the dominant-sign choice on a hand-built column pair.
"""
import math
import warnings

import numpy as np

__all__ = [
    "check_random_state", "norm", "unfold", "fold", "khatri_rao",
    "unfolding_dot_khatri_rao", "unfolding_dot_khatri_rao_memory",
    "unfolding_dot_khatri_rao_einsum", "unfolding_dot_khatri_rao_coo", "mode_dot", "multi_mode_dot",
    "validate_cp_tensor", "validate_cp_rank", "cp_normalize", "cp_to_tensor",
    "cp_norm", "cp_lstsq_grad", "random_cp", "svd_flip", "truncated_svd", "symeig_svd",
    "svd_interface", "make_svd_non_negative", "initialize_cp", "error_calc", "parafac",
    "non_negative_parafac", "hals_nnls", "non_negative_parafac_hals",
    "validate_tucker_rank", "tucker_to_tensor", "initialize_tucker",
    "partial_tucker", "tucker",
]


# --------------------------------------------------------------------------
# L1 primitives
# --------------------------------------------------------------------------
def check_random_state(seed):
    """``backend/core.py:162-188``: None -> numpy's global RandomState,
    int -> fresh ``RandomState(seed)``, RandomState -> itself."""
    if seed is None:
        return np.random.mtrand._rand
    if isinstance(seed, int):
        return np.random.RandomState(seed)
    if isinstance(seed, np.random.RandomState):
        return seed
    raise ValueError("Seed should be None, int or np.random.RandomState")


def norm(tensor, order=2, axis=None):
    """``backend/core.py:714-739``: l-``order`` norm, ``sqrt(sum(|x|^2))``
    for order 2."""
    if axis == ():
        axis = None
    a = np.abs(tensor)
    if order == "inf":
        return np.max(a, axis=axis)
    if order == 1:
        return np.sum(a, axis=axis)
    if order == 2:
        return np.sqrt(np.sum(a ** 2, axis=axis))
    return np.sum(a ** order, axis=axis) ** (1 / order)


# --------------------------------------------------------------------------
# L2 layout: unfold / fold  (base.py:39-79)
# --------------------------------------------------------------------------
def unfold(tensor, mode):
    """``base.py:39-53``: move ``mode`` to the front, then C-order reshape
    to ``(I_mode, -1)``.  Columns run row-major over the remaining modes in
    increasing order."""
    t = np.moveaxis(tensor, mode, 0)
    return np.reshape(t, (tensor.shape[mode], -1))


def fold(unfolded, mode, shape):
    """``base.py:56-79``: inverse of :func:`unfold`."""
    full = list(shape)
    lead = full.pop(mode)
    full.insert(0, lead)
    return np.moveaxis(np.reshape(unfolded, full), 0, mode)


# --------------------------------------------------------------------------
# L3 tensor algebra
# --------------------------------------------------------------------------
def khatri_rao(matrices, weights=None, skip_matrix=None, mask=None):
    """``tenalg/core_tenalg/_khatri_rao.py:9-109``.

    Column-wise Kronecker product, first matrix varying slowest.  ``weights``
    scale the first (non-skipped) matrix (``:95-99``); one matrix is
    returned unchanged (``:68-69``, weights ignored there, as in the
    reference); 1-D inputs are treated as single columns (``:71-79``)."""
    if skip_matrix is not None:
        matrices = [m for i, m in enumerate(matrices) if i != skip_matrix]
    if len(matrices) == 1:
        return matrices[0]
    if np.ndim(matrices[0]) == 2:
        ncol = matrices[0].shape[1]
    else:
        ncol = 1
        matrices = [np.reshape(m, (-1, 1)) for m in matrices]
        warnings.warn("Khatri-rao of a series of vectors instead of matrices. "
                      "Considering each as a matrix with 1 column.")
    for i, m in enumerate(matrices):
        if np.ndim(m) != 2:
            raise ValueError("All the matrices must have exactly 2 dimensions!"
                             f"Matrix {i} has dimension {np.ndim(m)} != 2.")
        if m.shape[1] != ncol:
            raise ValueError("All matrices must have same number of columns!"
                             f"Matrix {i} has {m.shape[1]} columns != {ncol}.")
    acc = matrices[0] if weights is None else matrices[0] * np.reshape(weights, (1, -1))
    for m in matrices[1:]:
        acc = np.reshape(acc[:, None, :] * m[None, :, :], (-1, ncol))
    if mask is not None:
        acc = acc * np.reshape(mask, (-1, 1))
    return acc


def unfolding_dot_khatri_rao(tensor, cp_tensor, mode):
    """``tenalg/core_tenalg/mttkrp.py:47-50``: MTTKRP as
    ``unfold(X, mode) @ conj(khatri_rao(factors, weights, skip=mode))``
    (KR materialised, one BLAS GEMM)."""
    weights, factors = cp_tensor
    kr = khatri_rao(factors, weights=weights, skip_matrix=mode)
    return np.dot(unfold(tensor, mode), np.conj(kr))


def unfolding_dot_khatri_rao_memory(tensor, cp_tensor, mode):
    """``tenalg/core_tenalg/mttkrp.py:89-104``: one tensor-times-vector
    chain per rank-one component, weights applied at the end."""
    weights, factors = cp_tensor
    rank = factors[0].shape[1]
    cols = []
    for r in range(rank):
        cols.append(multi_mode_dot(tensor, [np.conj(f[:, r]) for f in factors], skip=mode))
    out = np.stack(cols, axis=1)
    if weights is not None:
        out = out * np.reshape(weights, (1, -1))
    return out


def unfolding_dot_khatri_rao_einsum(tensor, cp_tensor, mode):
    """``tenalg/einsum_tenalg/mttkrp.py:6-39``: the same contraction as a
    single einsum, ``weights=None`` treated as ones (``:26-27``)."""
    weights, factors = cp_tensor
    rank = factors[0].shape[1]
    if weights is None:
        weights = np.ones(rank, dtype=factors[0].dtype)
    letters = "abcdefghijklmnopqstuvwxyz"
    n = tensor.ndim
    t_sub = letters[:n]
    ops = [tensor, weights]
    subs = [t_sub, "r"]
    for i, f in enumerate(factors):
        if i != mode:
            ops.append(np.conj(f))
            subs.append(letters[i] + "r")
    spec = ",".join(subs) + "->" + letters[mode] + "r"
    return np.einsum(spec, *ops)


def unfolding_dot_khatri_rao_coo(coords, values, shape, cp_tensor, mode):
    """MTTKRP of a sparse tensor in coordinate form (``contrib/sparse/
    cp_tensor.py:9``: ``unfolding_dot_khatri_rao`` over a sparse backend):
    each nonzero adds ``val * w * prod_{m != mode} A_m[i_m, :]`` to row
    ``i_mode`` (repeated coordinates add up, as when densified).  At order 2
    the weights are not applied: the sparse backend wraps the core
    ``unfolding_dot_khatri_rao`` (``contrib/sparse/cp_tensor.py:9``), whose
    Khatri-Rao product of a single matrix drops them (``_khatri_rao.py:67-69``)."""
    weights, factors = cp_tensor
    rank = factors[0].shape[1]
    contrib = np.asarray(values, dtype=np.float64)[:, None] * np.ones((1, rank))
    for m in range(len(shape)):
        if m != mode:
            contrib = contrib * factors[m][coords[m]]
    if weights is not None and len(shape) > 2:
        contrib = contrib * weights[None, :]
    out = np.zeros((shape[mode], rank))
    np.add.at(out, coords[mode], contrib)
    return out


def mode_dot(tensor, matrix_or_vector, mode, transpose=False):
    """``tenalg/core_tenalg/n_mode_product.py:5-76``: ``fold(U @ unfold(X,
    mode))``; a vector removes the mode; ``transpose`` uses the conjugate
    transpose; ValueError on any shape mismatch."""
    new_shape = list(tensor.shape)
    mv = matrix_or_vector
    if np.ndim(mv) == 2:
        dim = 0 if transpose else 1
        if mv.shape[dim] != tensor.shape[mode]:
            raise ValueError(
                f"shapes {tensor.shape} and {mv.shape} not aligned in mode-{mode} multiplication: "
                f"{tensor.shape[mode]} (mode {mode}) != {mv.shape[dim]} (dim 1 of matrix)")
        if transpose:
            mv = np.conj(np.transpose(mv))
        new_shape[mode] = mv.shape[0]
        is_vec = False
    elif np.ndim(mv) == 1:
        if mv.shape[0] != tensor.shape[mode]:
            raise ValueError(
                f"shapes {tensor.shape} and {mv.shape} not aligned for mode-{mode} multiplication: "
                f"{tensor.shape[mode]} (mode {mode}) != {mv.shape[0]} (vector size)")
        if len(new_shape) > 1:
            new_shape.pop(mode)
        else:
            new_shape = ()
        is_vec = True
    else:
        raise ValueError("Can only take n_mode_product with a vector or a matrix."
                         f"Provided array of dimension {np.ndim(mv)} not in [1, 2].")
    res = np.dot(mv, unfold(tensor, mode))
    if is_vec:
        return np.reshape(res, new_shape)
    return fold(res, mode, new_shape)


def multi_mode_dot(tensor, matrix_or_vec_list, modes=None, skip=None, transpose=False):
    """``tenalg/core_tenalg/n_mode_product.py:79-135``: chain of
    :func:`mode_dot` sorted by mode (stable); ``skip`` is a position in the
    sorted list; each vector contraction shifts later modes down by one."""
    if modes is None:
        modes = range(len(matrix_or_vec_list))
    pairs = sorted(zip(matrix_or_vec_list, modes), key=lambda p: p[1])
    res = tensor
    removed = 0
    for pos, (mv, mode) in enumerate(pairs):
        if skip is not None and pos == skip:
            continue
        if transpose:
            res = mode_dot(res, np.conj(np.transpose(mv)), mode - removed)
        else:
            res = mode_dot(res, mv, mode - removed)
        if np.ndim(mv) == 1:
            removed += 1
    return res


# --------------------------------------------------------------------------
# CP tensors  (cp_tensor.py)
# --------------------------------------------------------------------------
def validate_cp_tensor(cp_tensor):
    """``cp_tensor.py:163-214``: returns (shape, rank); ValueError on
    inconsistent factor widths or weights length."""
    weights, factors = cp_tensor
    if np.ndim(factors[0]) == 2:
        rank = int(factors[0].shape[1])
    elif np.ndim(factors[0]) == 1:
        rank = 1
    else:
        raise ValueError("Got a factor with 3 dimensions but CP factors should be at most 2D, "
                         "of shape (size, rank).")
    shape = []
    for i, f in enumerate(factors):
        s = f.shape
        size, r = (s if len(s) == 2 else (s[0], 1))
        if r != rank:
            raise ValueError("All the factors of a CP tensor should have the same number of column."
                             f"However, factors[0].shape[1]={rank} but factors[{i}].shape[1]={r}.")
        shape.append(size)
    if weights is not None and np.shape(weights) != (rank,):
        raise ValueError(f"Given factors for a rank-{rank} CP tensor but len(weights)={np.shape(weights)}.")
    return tuple(shape), rank


def validate_cp_rank(tensor_shape, rank="same", rounding="round"):
    """``cp_tensor.py:240-275``."""
    fun = {"ceil": np.ceil, "floor": np.floor, "round": np.round}.get(rounding)
    if fun is None:
        raise ValueError(f"Rounding should be of round, floor or ceil, but got {rounding}")
    if rank == "same":
        rank = 1.0
    if isinstance(rank, float):
        rank = int(fun(np.prod(tensor_shape) * rank / np.sum(tensor_shape)))
    return rank


def cp_normalize(cp_tensor):
    """``cp_tensor.py:278-321``: weights folded into the first factor, then
    every factor's columns scaled to unit l2 norm (zero columns kept)."""
    _, rank = validate_cp_tensor(cp_tensor)
    weights, factors = cp_tensor
    if weights is None:
        weights = np.ones(rank, dtype=factors[0].dtype)
    out = []
    for i, f in enumerate(factors):
        if i == 0:
            f = f * weights
            weights = np.ones(rank, dtype=f.dtype)
        scales = norm(f, axis=0)
        safe = np.where(scales == 0, np.ones_like(scales), scales)
        weights = weights * scales
        out.append(f / np.reshape(safe, (1, -1)))
    return weights, out


def cp_to_tensor(cp_tensor, mask=None):
    """``cp_tensor.py:433-485``: ``fold((A_0 * w) @ khatri_rao(A_1..)^T, 0)``,
    or the masked row-sum form."""
    shape, _ = validate_cp_tensor(cp_tensor)
    weights, factors = cp_tensor
    if len(shape) == 1:
        return np.sum(weights * factors[0], axis=1)
    if weights is None:
        weights = 1
    if mask is None:
        full = np.dot(factors[0] * weights, np.transpose(khatri_rao(factors, skip_matrix=0)))
    else:
        full = np.sum(khatri_rao([factors[0] * weights] + list(factors[1:]), mask=mask), axis=1)
    return fold(full, 0, shape)


def cp_norm(cp_tensor):
    """``cp_tensor.py:614-644``: ``sqrt(sum((*_m A_m^T A_m) * w w^T))``."""
    validate_cp_tensor(cp_tensor)
    weights, factors = cp_tensor
    r = factors[0].shape[1]
    acc = np.ones((r, r), dtype=factors[0].dtype)
    for f in factors:
        acc = acc * np.dot(np.transpose(f), np.conj(f))
    if weights is not None:
        acc = acc * (np.reshape(weights, (-1, 1)) * np.reshape(weights, (1, -1)))
    return np.sqrt(np.sum(acc))


def cp_lstsq_grad(cp_tensor, tensor, return_loss=False):
    """``cp_tensor.py:377-430`` (no mask): the residual ``X - [[w; A]]``, then
    ``grad_n = -unfolding_dot_khatri_rao(residual, (w, A), n)`` per mode and
    the loss ``0.5 * sum(residual ** 2)``."""
    weights, factors = cp_tensor
    diff = tensor - cp_to_tensor(cp_tensor)
    grads = [-unfolding_dot_khatri_rao(diff, cp_tensor, n) for n in range(len(factors))]
    if return_loss:
        return (None, grads), 0.5 * np.sum(diff ** 2)
    return None, grads


def random_cp(shape, rank, full=False, orthogonal=False, random_state=None,
              normalise_factors=True, dtype=np.float64):
    """``random/base.py:64-114``: one ``random_sample((I_m, R))`` draw per mode
    in mode order, then ones weights; optional QR, full or normalised."""
    rank = validate_cp_rank(shape, rank)
    rng = check_random_state(random_state)
    factors = [np.asarray(rng.random_sample((s, rank)), dtype=dtype) for s in shape]
    weights = np.ones(rank, dtype=dtype)
    if orthogonal:
        factors = [np.linalg.qr(f)[0] for f in factors]
    if full:
        return cp_to_tensor((weights, factors))
    if normalise_factors:
        return cp_normalize((weights, factors))
    return weights, factors


# --------------------------------------------------------------------------
# SVD interface  (tenalg/svd.py)
# --------------------------------------------------------------------------
def svd_flip(U, V, u_based_decision=True):
    """``tenalg/svd.py:13-65``: make the largest-|.| entry of every U column
    (first one on ties, np.argmax) positive; V rows follow."""
    if u_based_decision:
        idx = np.argmax(np.abs(U), axis=0)
        signs = np.sign(U[idx, np.arange(U.shape[1])])
        U = U * np.conj(signs)
        if V.shape[0] > U.shape[1]:
            signs = np.concatenate((signs, np.ones(V.shape[0] - U.shape[1], dtype=V.dtype)))
        V = V * signs[: V.shape[0]][:, None]
    else:
        idx = np.argmax(np.abs(V), axis=1)
        signs = np.sign(V[np.arange(V.shape[0]), idx])
        V = V * np.conj(signs[:, None])
        if U.shape[1] > V.shape[0]:
            signs = np.concatenate((signs, np.ones(U.shape[1] - V.shape[0], dtype=V.dtype)))
        U = U * signs[: U.shape[1]]
    return U, V


def _svd_checks(matrix, n_eigenvecs=None):
    """``tenalg/svd.py:173-208``."""
    if np.ndim(matrix) != 2:
        raise ValueError(f"matrix be a matrix. matrix.ndim is {np.ndim(matrix)} != 2")
    d1, d2 = matrix.shape
    lo, hi = min(d1, d2), max(d1, d2)
    if n_eigenvecs is None:
        n_eigenvecs = hi
    if n_eigenvecs > hi:
        warnings.warn(f"Trying to compute SVD with n_eigenvecs={n_eigenvecs}, which is larger "
                      f"than max(matrix.shape)={hi}. Setting n_eigenvecs to {hi}.")
        n_eigenvecs = hi
    return n_eigenvecs, lo, hi


def truncated_svd(matrix, n_eigenvecs=None, **kwargs):
    """``tenalg/svd.py:211-235``: LAPACK SVD (full matrices only when more
    vectors than min(shape) are asked for), truncated."""
    n, lo, _ = _svd_checks(matrix, n_eigenvecs)
    U, S, V = np.linalg.svd(matrix, full_matrices=n > lo)
    return U[:, :n], S[:n], V[:n, :]


def symeig_svd(matrix, n_eigenvecs=None, **kwargs):
    """``tenalg/svd.py:238-285``: SVD through ``eigh`` of the smaller Gram,
    eigenvalues clipped at eps, then flipped to descending order."""
    n, _, _ = _svd_checks(matrix, n_eigenvecs)
    d1, d2 = matrix.shape
    if d1 > d2:
        S, U = np.linalg.eigh(np.dot(matrix, np.transpose(matrix)))
        S = np.sqrt(np.clip(S, np.finfo(S.dtype).eps, None))
        V = np.dot(np.transpose(matrix), U / np.reshape(S, (1, -1)))
    else:
        S, V = np.linalg.eigh(np.dot(np.transpose(matrix), matrix))
        S = np.sqrt(np.clip(S, np.finfo(S.dtype).eps, None))
        U = np.dot(matrix, V) / np.reshape(S, (1, -1))
    U, S, V = np.flip(U, axis=1), np.flip(S), np.flip(np.transpose(V), axis=0)
    return U[:, : min(d1, n)], S[: min(d1, d2, n)], V[: min(d2, n), :]


def make_svd_non_negative(tensor, U, S, V, nntype=True):
    """``tenalg/svd.py:68-127``: NNDSVD (``'nndsvd'``: small entries set to
    zero by soft thresholding at eps) or NNDSVDa (``True`` / ``'nndsvda'``:
    small entries set to the mean of ``tensor``)."""
    if nntype is True:
        nntype = "nndsvda"
    W = np.zeros_like(U)
    H = np.zeros_like(V)
    W[:, 0] = np.sqrt(S[0]) * np.abs(U[:, 0])
    H[0, :] = np.sqrt(S[0]) * np.abs(V[0, :])
    for j in range(1, min(U.shape[1], V.shape[0])):
        x, y = U[:, j], V[j, :]
        x_p, y_p = np.clip(x, 0.0, None), np.clip(y, 0.0, None)
        x_n, y_n = np.abs(np.clip(x, None, 0.0)), np.abs(np.clip(y, None, 0.0))
        x_p_nrm, y_p_nrm = np.linalg.norm(x_p), np.linalg.norm(y_p)
        x_n_nrm, y_n_nrm = np.linalg.norm(x_n), np.linalg.norm(y_n)
        m_p, m_n = x_p_nrm * y_p_nrm, x_n_nrm * y_n_nrm
        if m_p > m_n:
            u, v, sigma = x_p / x_p_nrm, y_p / y_p_nrm, m_p
        else:
            u, v, sigma = x_n / x_n_nrm, y_n / y_n_nrm, m_n
        lbd = np.sqrt(S[j] * sigma)
        W[:, j] = lbd * u
        H[j, :] = lbd * v
    eps = np.finfo(tensor.dtype).eps
    if nntype == "nndsvd":
        W = np.sign(W) * np.clip(np.abs(W) - eps, 0.0, None)
        H = np.sign(H) * np.clip(np.abs(H) - eps, 0.0, None)
    elif nntype == "nndsvda":
        avg = np.mean(tensor)
        W = np.where(W < eps, np.ones(W.shape, dtype=W.dtype) * avg, W)
        H = np.where(H < eps, np.ones(H.shape, dtype=H.dtype) * avg, H)
    else:
        raise ValueError(f'Invalid nntype parameter: got {nntype} instead of one of ("nndsvd", "nndsvda")')
    return W, H


def svd_interface(matrix, method="truncated_svd", n_eigenvecs=None, flip_sign=True,
                  u_based_flip_sign=True, non_negative=None, mask=None, n_iter_mask_imputation=5, **kwargs):
    """``tenalg/svd.py:366-447``.  With ``mask`` (and ``n_eigenvecs``) the
    missing entries are filled ``n_iter_mask_imputation`` times from the
    current truncated SVD, U diag(S) V, and the SVD recomputed (``:431-439``)."""
    if method == "truncated_svd":
        fun = truncated_svd
    elif method == "symeig_svd":
        fun = symeig_svd
    elif callable(method):
        fun = method
    else:
        raise ValueError(f"Got svd={method}. However, the possible choices are "
                         "['truncated_svd', 'symeig_svd', 'randomized_svd'] or to pass a callable.")
    U, S, V = fun(matrix, n_eigenvecs=n_eigenvecs)
    if mask is not None and n_eigenvecs is not None:
        for _ in range(n_iter_mask_imputation):
            St = np.eye(U.shape[1], V.shape[0], dtype=matrix.dtype)
            St[np.arange(S.shape[0]), np.arange(S.shape[0])] = S
            matrix = matrix * mask + (U @ St @ V) * (1 - mask)
            U, S, V = fun(matrix, n_eigenvecs=n_eigenvecs)
    if flip_sign:
        U, V = svd_flip(U, V, u_based_decision=u_based_flip_sign)
    if non_negative is not False and non_negative is not None:
        U, V = make_svd_non_negative(matrix, U, S, V, non_negative)
    return U, S, V


# --------------------------------------------------------------------------
# CP-ALS  (decomposition/_cp.py)
# --------------------------------------------------------------------------
def initialize_cp(tensor, rank, init="svd", svd="truncated_svd", random_state=None,
                  normalize_factors=False, non_negative=False, mask=None, svd_mask_repeats=5):
    """``decomposition/_cp.py:26-137``; with ``mask`` the SVD init imputes the
    unfolded mask's missing entries (``:66-76``)."""
    rng = check_random_state(random_state)
    if isinstance(init, str) and init == "random":
        weights, factors = random_cp(tensor.shape, rank, normalise_factors=False,
                                     random_state=rng, dtype=tensor.dtype)
    elif isinstance(init, str) and init == "svd":
        factors = []
        for mode in range(tensor.ndim):
            U, S, _ = svd_interface(unfold(tensor, mode), n_eigenvecs=rank, method=svd,
                                    non_negative=non_negative,
                                    mask=None if mask is None else unfold(mask, mode),
                                    n_iter_mask_imputation=svd_mask_repeats)
            if mode == 0:
                k = min(rank, S.shape[0])
                U = U.copy()
                U[:, :k] = U[:, :k] * S[:k]
            if tensor.shape[mode] < rank:
                pad = np.asarray(rng.random_sample((U.shape[0], rank - tensor.shape[mode])),
                                 dtype=tensor.dtype)
                U = np.concatenate([U, pad], axis=1)
            factors.append(U[:, :rank])
        weights = np.ones(rank, dtype=tensor.dtype)
    elif isinstance(init, (tuple, list)):
        weights, factors = init
        factors = [np.array(f) for f in factors]
        if weights is None or np.all(weights == 1):
            return np.ones(rank, dtype=factors[0].dtype), factors
        avg = np.prod(weights) ** (1.0 / weights.shape[0])
        return np.ones(rank, dtype=factors[0].dtype), [f * avg for f in factors]
    else:
        raise ValueError(f'Initialization method "{init}" not recognized')
    if non_negative:
        factors = [np.abs(f) for f in factors]
    if normalize_factors:
        weights, factors = cp_normalize((weights, factors))
    return weights, factors


def error_calc(tensor, norm_tensor, weights, factors, mttkrp=None, mask=None):
    """``decomposition/_cp.py:162-227`` (no sparsity).

    With the last mode's MTTKRP: ``sqrt(|‖X‖² + ‖M‖² − 2<X, M>|)`` where
    ``<X, M> = sum(mttkrp * A_last)``; otherwise the full reconstruction.
    With ``mask`` the tensor's missing entries are first replaced by the
    model and its norm recomputed; returns ``(error, tensor, norm_tensor)``
    then."""
    if mask is not None:
        rec = cp_to_tensor((weights, factors))
        tensor = tensor * mask + rec * (1 - mask)
        norm_tensor = norm(tensor, 2)
        return norm(tensor - rec, 2), tensor, norm_tensor
    if mttkrp is None:
        rec = cp_to_tensor((weights, factors))
        return norm(tensor - rec, 2)
    fn = cp_norm((weights, factors))
    iprod = np.sum(np.sum(mttkrp * np.conj(factors[-1]), axis=0))
    return np.sqrt(np.abs(norm_tensor ** 2 + fn ** 2 - 2 * iprod))


def parafac(tensor, rank, n_iter_max=100, init="svd", svd="truncated_svd",
            normalize_factors=False, orthogonalise=False, tol=1e-8, random_state=None,
            verbose=0, return_errors=False, l2_reg=0, cvg_criterion="abs_rec_error",
            fixed_modes=None, linesearch=False, callback=None, mttkrp_fn=None,
            trace=None, mask=None, svd_mask_repeats=5):
    """``decomposition/_cp.py:230-543`` (no sparsity).  With ``mask`` the
    working tensor's missing entries are re-imputed from the model after
    every sweep (``:437-445``) and the error is that of the imputed tensor.

    ``mttkrp_fn`` swaps the MTTKRP (default: the materialised-KR form of
    ``mttkrp.py:47-50``); ``trace``, if a list, receives per-iteration
    records (line-search jump / accepted) for the tests."""
    mttkrp_fn = mttkrp_fn or unfolding_dot_khatri_rao
    rank = validate_cp_rank(tensor.shape, rank=rank)
    if orthogonalise and not isinstance(orthogonalise, int):
        orthogonalise = n_iter_max
    if linesearch:
        acc_pow, acc_fail, max_fail = 2.0, 0, 4
    weights, factors = initialize_cp(tensor, rank, init=init, svd=svd,
                                     random_state=random_state,
                                     normalize_factors=normalize_factors, mask=mask,
                                     svd_mask_repeats=svd_mask_repeats)
    factors = list(factors)
    rec_errors = []
    norm_tensor = norm(tensor, 2)
    Id = np.eye(rank, dtype=tensor.dtype) * l2_reg if l2_reg else 0
    fixed_modes = list(fixed_modes) if fixed_modes is not None else []
    if fixed_modes == list(range(tensor.ndim)):
        return (weights, factors)
    if tensor.ndim - 1 in fixed_modes:
        warnings.warn("You asked for fixing the last mode, which is not supported.")
        fixed_modes.remove(tensor.ndim - 1)
    modes_list = [m for m in range(tensor.ndim) if m not in fixed_modes]

    if callback is not None:
        if mask is not None:
            err0, _, norm_tensor = error_calc(tensor, norm_tensor, weights, factors, mask=mask)
        else:
            err0 = error_calc(tensor, norm_tensor, weights, factors)
        callback((weights, factors), err0 / norm_tensor)

    rec_error = None
    for iteration in range(n_iter_max):
        if orthogonalise and iteration <= orthogonalise:
            factors = [np.linalg.qr(f)[0] if min(f.shape) >= rank else f for f in factors]
        if linesearch and iteration % 2 == 0:
            factors_last = [f.copy() for f in factors]
            weights_last = weights.copy()
        for mode in modes_list:
            V = np.ones((rank, rank), dtype=tensor.dtype)
            for i, f in enumerate(factors):
                if i != mode:
                    V = V * np.dot(np.conj(np.transpose(f)), f)
            V = V + Id
            V = np.reshape(weights, (-1, 1)) * V * np.reshape(weights, (1, -1))
            mttkrp = mttkrp_fn(tensor, (weights, factors), mode)
            factors[mode] = np.transpose(np.linalg.solve(np.conj(np.transpose(V)),
                                                         np.transpose(mttkrp)))
        line_iter = bool(linesearch and iteration % 2 == 0 and iteration > 5)
        if (tol or return_errors) and not line_iter:
            if mask is not None:
                unnorm, tensor, norm_tensor = error_calc(tensor, norm_tensor, weights, factors, mask=mask)
            else:
                unnorm = error_calc(tensor, norm_tensor, weights, factors, mttkrp)
        elif mask is not None:
            tensor = tensor * mask + cp_to_tensor((weights, factors)) * (1 - mask)
        rec = {"iteration": iteration}
        if line_iter:
            jump = iteration ** (1.0 / acc_pow)
            new_w = weights_last + (weights - weights_last) * jump
            new_f = [factors_last[i] + (factors[i] - factors_last[i]) * jump
                     for i in range(tensor.ndim)]
            if mask is not None:
                new_err, new_tensor, new_norm = error_calc(tensor, norm_tensor, new_w, new_f, mask=mask)
            else:
                new_err = error_calc(tensor, norm_tensor, new_w, new_f)
                new_tensor, new_norm = tensor, norm_tensor
            rec["jump"] = jump
            rec["new_error"] = new_err / new_norm
            rec["last_error"] = rec_errors[-1]
            if new_err / new_norm < rec_errors[-1]:
                factors, weights = new_f, new_w
                tensor, norm_tensor = new_tensor, new_norm
                unnorm = new_err
                acc_fail = 0
                rec["accepted"] = True
            else:
                if mask is not None:
                    unnorm, tensor, norm_tensor = error_calc(tensor, norm_tensor, weights, factors, mask=mask)
                else:
                    unnorm = error_calc(tensor, norm_tensor, weights, factors, mttkrp)
                acc_fail += 1
                rec["accepted"] = False
                if acc_fail == max_fail:
                    acc_pow += 1.0
                    acc_fail = 0
        if trace is not None:
            trace.append(rec)
        if (tol or return_errors) and not line_iter:
            rec_error = unnorm / norm_tensor
            rec_errors.append(rec_error)
        if callback is not None:
            if callback((weights, factors), rec_error) is True:
                break
        if tol:
            if iteration >= 1:
                dec = rec_errors[-2] - rec_errors[-1]
                if cvg_criterion == "abs_rec_error":
                    stop = abs(dec) < tol
                elif cvg_criterion == "rec_error":
                    stop = dec < tol
                else:
                    raise TypeError("Unknown convergence criterion")
                if stop:
                    break
        if normalize_factors:
            weights, factors = cp_normalize((weights, factors))
    if return_errors:
        return (weights, factors), rec_errors
    return (weights, factors)


# --------------------------------------------------------------------------
# Tucker  (tucker_tensor.py, decomposition/_tucker.py)
# --------------------------------------------------------------------------
# Non-negative CP by multiplicative updates  (decomposition/_nn_cp.py)
# --------------------------------------------------------------------------
def non_negative_parafac(tensor, rank, n_iter_max=100, init="svd", svd="truncated_svd", tol=10e-7,
                         random_state=None, normalize_factors=False, return_errors=False,
                         cvg_criterion="abs_rec_error", fixed_modes=None):
    """``decomposition/_nn_cp.py:26-180`` (no mask, no verbose output).

    Per mode (``:111-135``): ``accum`` = Hadamard of the other modes' Grams
    times ``w w^T``; ``A <- A * clip(M, eps) / clip(A accum, eps)`` with
    ``eps`` the machine epsilon of the tensor's dtype.  The error (``:138-150``)
    comes from the last mode's MTTKRP as in CP-ALS."""
    epsilon = np.finfo(tensor.dtype).eps
    rank = validate_cp_rank(tensor.shape, rank=rank)
    weights, factors = initialize_cp(tensor, rank, init=init, svd=svd, non_negative=True,
                                     random_state=random_state, normalize_factors=normalize_factors)
    factors = [np.array(f) for f in factors]
    rec_errors = []
    norm_tensor = norm(tensor, 2)
    fixed_modes = [] if fixed_modes is None else list(fixed_modes)
    if tensor.ndim - 1 in fixed_modes:
        warnings.warn("You asked for fixing the last mode, which is not supported while tol is fixed.")
        fixed_modes.remove(tensor.ndim - 1)
    modes_list = [mode for mode in range(tensor.ndim) if mode not in fixed_modes]
    mttkrp = None
    for iteration in range(n_iter_max):
        for mode in modes_list:
            accum = None
            for e in [i for i in range(len(factors)) if i != mode]:
                g = np.dot(factors[e].T, factors[e])
                accum = g if accum is None else accum * g
            accum = weights.reshape(-1, 1) * accum * weights.reshape(1, -1)
            mttkrp = unfolding_dot_khatri_rao(tensor, (weights, factors), mode)
            numerator = np.clip(mttkrp, epsilon, None)
            denominator = np.clip(np.dot(factors[mode], accum), epsilon, None)
            factors[mode] = factors[mode] * numerator / denominator
            if normalize_factors and mode != modes_list[-1]:
                weights, factors = cp_normalize((weights, factors))
        if tol:
            rec_error = error_calc(tensor, norm_tensor, weights, factors, mttkrp=mttkrp) / norm_tensor
            rec_errors.append(rec_error)
            if iteration >= 1:
                dec = rec_errors[-2] - rec_errors[-1]
                if cvg_criterion == "abs_rec_error":
                    stop = abs(dec) < tol
                elif cvg_criterion == "rec_error":
                    stop = dec < tol
                else:
                    raise TypeError("Unknown convergence criterion")
                if stop:
                    break
        if normalize_factors:
            weights, factors = cp_normalize((weights, factors))
    if return_errors:
        return (weights, factors), rec_errors
    return weights, factors


def hals_nnls(UtM, UtU, V=None, n_iter_max=500, tol=1e-8, sparsity_coefficient=None, ridge_coefficient=None,
              nonzero_rows=False, exact=False, epsilon=0.0):
    """``solvers/nnls.py:5-173``: block-coordinate (HALS) descent for
    min_{V >= 0} ||M - U V||^2 given UtM = U^T M (rank x n) and UtU = U^T U;
    rows of V are updated in order, each with the rows before it already
    updated, until the iteration's sum of ||V - newV||^2 (the new row
    broadcast against every row of the current V, ``nnls.py:147``) falls
    below ``tol`` times the first iteration's."""
    rank = UtM.shape[0]
    if V is None:
        V = np.linalg.solve(UtU, UtM)
        V = np.clip(V, 0.0, None)
        scale = np.sum(UtM * V) / np.sum(UtU * np.dot(V, V.T))
        V = V * scale
    V = np.array(V)
    if exact:
        n_iter_max = 50000
        tol = 1e-16
    rec_error0 = None
    for iteration in range(n_iter_max):
        rec_error = 0.0
        for k in range(rank):
            if UtU[k, k]:
                num = UtM[k, :] - np.dot(UtU[k, :], V) + UtU[k, k] * V[k, :]
                den = UtU[k, k]
                if sparsity_coefficient is not None:
                    num = num - sparsity_coefficient
                if ridge_coefficient is not None:
                    den = den + 2 * ridge_coefficient
                newV = np.clip(num / den, epsilon, None)
                # the reference subtracts the new row from every row of V
                # (broadcast), not from row k alone: the stopping rule sums
                # ||V[j, :] - newV||^2 over all rows j
                rec_error += np.linalg.norm(V - newV) ** 2
                V[k, :] = newV
                if nonzero_rows and np.all(V[k, :] == 0):
                    V[k, :] = np.finfo(V.dtype).eps * np.max(V)
            elif nonzero_rows:
                raise ValueError("Column " + str(k) + " of U is zero with nonzero condition")
        if iteration == 0:
            rec_error0 = rec_error
        if rec_error < tol * rec_error0:
            break
    return V


def non_negative_parafac_hals(tensor, rank, n_iter_max=100, init="svd", svd="truncated_svd", tol=10e-8,
                              random_state=None, sparsity_coefficients=None, fixed_modes=None, nn_modes="all",
                              exact=False, normalize_factors=False, return_errors=False,
                              cvg_criterion="abs_rec_error"):
    """``decomposition/_nn_cp.py:183-360`` (no verbose output): per mode the
    Hadamard of the other Grams times w w^T, the MTTKRP, then
    ``hals_nnls(MTTKRP^T, pseudo_inverse, A^T, n_iter_max=100)`` for
    non-negative modes and a linear solve for the others; the error from the
    last mode's MTTKRP."""
    weights, factors = initialize_cp(tensor, rank, init=init, svd=svd, non_negative=True,
                                     random_state=random_state, normalize_factors=normalize_factors)
    factors = [np.array(f) for f in factors]
    norm_tensor = norm(tensor, 2)
    n_modes = tensor.ndim
    if sparsity_coefficients is None or isinstance(sparsity_coefficients, float):
        sparsity_coefficients = [sparsity_coefficients] * n_modes
    sparsity_coefficients = list(sparsity_coefficients)
    fixed_modes = [] if fixed_modes is None else list(fixed_modes)
    if nn_modes == "all":
        nn_modes = set(range(n_modes))
    elif nn_modes is None:
        nn_modes = set()
    for fixed_value in fixed_modes:
        sparsity_coefficients[fixed_value] = None
    modes = [mode for mode in range(n_modes) if mode not in fixed_modes]
    rec_errors = []
    mttkrp = None
    for iteration in range(n_iter_max):
        for mode in modes:
            pseudo_inverse = np.ones((rank, rank), dtype=tensor.dtype)
            for i, factor in enumerate(factors):
                if i != mode:
                    pseudo_inverse = pseudo_inverse * np.dot(factor.T, factor)
            pseudo_inverse = weights.reshape(-1, 1) * pseudo_inverse * weights.reshape(1, -1)
            mttkrp = unfolding_dot_khatri_rao(tensor, (weights, factors), mode)
            if mode in nn_modes:
                nn_factor = hals_nnls(mttkrp.T, pseudo_inverse, factors[mode].T, n_iter_max=100,
                                      sparsity_coefficient=sparsity_coefficients[mode], exact=exact)
                factors[mode] = nn_factor.T
            else:
                factors[mode] = np.linalg.solve(pseudo_inverse.T, mttkrp.T).T
            if normalize_factors and mode != modes[-1]:
                weights, factors = cp_normalize((weights, factors))
        if tol:
            factors_norm = cp_norm((weights, factors))
            iprod = np.sum(np.sum(mttkrp * factors[-1], axis=0))
            rec_error = np.sqrt(np.abs(norm_tensor ** 2 + factors_norm ** 2 - 2 * iprod)) / norm_tensor
            rec_errors.append(rec_error)
            if iteration >= 1:
                dec = rec_errors[-2] - rec_errors[-1]
                if cvg_criterion == "abs_rec_error":
                    stop = abs(dec) < tol
                elif cvg_criterion == "rec_error":
                    stop = dec < tol
                else:
                    raise TypeError("Unknown convergence criterion")
                if stop:
                    break
        if normalize_factors:
            weights, factors = cp_normalize((weights, factors))
    if return_errors:
        return (weights, factors), rec_errors
    return weights, factors


# --------------------------------------------------------------------------
def validate_tucker_rank(tensor_shape, rank="same", rounding="round", fixed_modes=None):
    """``tucker_tensor.py:344-447`` (int and explicit-tuple forms, and the
    parameter-fraction form solved with brentq)."""
    fun = {"ceil": np.ceil, "floor": np.floor, "round": np.round}.get(rounding)
    if fun is None:
        raise ValueError(f"Rounding should be round, floor or ceil, but got {rounding}")
    if rank == "same":
        rank = 1.0
    if isinstance(rank, float):
        from scipy.optimize import brentq
        shape = list(tensor_shape)
        n_modes = len(shape)
        n_param = np.prod(shape)
        fixed = []
        n_fixed = 0
        if fixed_modes is not None:
            fixed = [(m, shape.pop(m)) for m in sorted(fixed_modes, reverse=True)][::-1]
            n_fixed = np.sum([s ** 2 for _, s in fixed])
            n_modes -= len(fixed)
        sq = np.sum([s ** 2 for s in shape])
        frac = brentq(lambda x: n_param * x ** n_modes + sq * x + n_fixed * x - rank * n_param,
                      0.0, max(rank, 1.0))
        out = [max(int(fun(s * frac)), 1) for s in shape]
        for m, s in fixed:
            out.insert(m, s)
        return out
    if isinstance(rank, int):
        if fixed_modes is None:
            return [rank] * len(tensor_shape)
        return [rank if i not in fixed_modes else s for i, s in enumerate(tensor_shape)]
    return rank


def tucker_to_tensor(tucker_tensor, skip_factor=None, transpose_factors=False, modes=None):
    """``tucker_tensor.py:50-75``."""
    core, factors = tucker_tensor
    return multi_mode_dot(core, factors, skip=skip_factor, transpose=transpose_factors,
                          modes=modes)


def initialize_tucker(tensor, rank, modes, random_state, init="svd", svd="truncated_svd"):
    """``decomposition/_tucker.py:25-102`` (no mask / non-negative)."""
    if isinstance(init, str) and init == "svd":
        factors = []
        for i, mode in enumerate(modes):
            U, _, _ = svd_interface(unfold(tensor, mode), n_eigenvecs=rank[i], method=svd)
            factors.append(U)
        core = multi_mode_dot(tensor, factors, modes=modes, transpose=True)
    elif isinstance(init, str) and init == "random":
        rng = check_random_state(random_state)
        core = np.asarray(rng.random_sample([rank[i] for i in range(len(modes))]) + 0.01,
                          dtype=tensor.dtype)
        factors = [np.asarray(rng.random_sample((tensor.shape[m], rank[i])), dtype=tensor.dtype)
                   for i, m in enumerate(modes)]
    else:
        core, factors = init
        factors = list(factors)
    return core, factors


def partial_tucker(tensor, rank, modes=None, n_iter_max=100, init="svd", tol=10e-5,
                   svd="truncated_svd", random_state=None, verbose=False):
    """``decomposition/_tucker.py:105-221``: HOOI.  Inside the loop the factor
    update always uses the default ``truncated_svd`` (``:197-202``)."""
    if modes is None:
        modes = list(range(tensor.ndim))
    if rank is None:
        rank = [tensor.shape[m] for m in modes]
    elif isinstance(rank, int):
        rank = tuple(rank for _ in modes)
    else:
        rank = tuple(rank)
    core, factors = initialize_tucker(tensor, rank, modes, init=init, svd=svd,
                                      random_state=random_state)
    rec_errors = []
    norm_tensor = norm(tensor, 2)
    for iteration in range(n_iter_max):
        for index, mode in enumerate(modes):
            y = multi_mode_dot(tensor, factors, modes=modes, skip=index, transpose=True)
            U, _, _ = svd_interface(unfold(y, mode), n_eigenvecs=rank[index])
            factors[index] = U
        core = multi_mode_dot(tensor, factors, modes=modes, transpose=True)
        rec_errors.append(math.sqrt(abs(norm_tensor ** 2 - norm(core, 2) ** 2)) / norm_tensor)
        if iteration > 1:
            if tol and abs(rec_errors[-2] - rec_errors[-1]) < tol:
                break
    return (core, factors), rec_errors


def tucker(tensor, rank, fixed_factors=None, n_iter_max=100, init="svd", return_errors=False,
           svd="truncated_svd", tol=10e-5, random_state=None, verbose=False):
    """``decomposition/_tucker.py:224-345``."""
    if fixed_factors:
        core, factors = init
        fixed_factors = sorted(fixed_factors)
        fixed = [(i, f) for i, f in enumerate(factors) if i in fixed_factors]
        modes_fixed = [i for i, _ in fixed]
        factors_fixed = [f for _, f in fixed]
        core = multi_mode_dot(core, factors_fixed, modes=modes_fixed)
        free = [(i, f) for i, f in enumerate(factors) if i not in fixed_factors]
        modes = [i for i, _ in free]
        (core, new_factors), errs = partial_tucker(
            tensor, rank=rank, modes=modes, n_iter_max=n_iter_max,
            init=(core, [f for _, f in free]), svd=svd, tol=tol, random_state=random_state)
        factors = list(new_factors)
        for i, e in enumerate(fixed_factors):
            factors.insert(e, factors_fixed[i])
        core = multi_mode_dot(core, factors_fixed, modes=modes_fixed, transpose=True)
        return (core, factors)
    modes = list(range(tensor.ndim))
    rank = validate_tucker_rank(tensor.shape, rank=rank)
    (core, factors), errs = partial_tucker(tensor, rank=rank, modes=modes,
                                           n_iter_max=n_iter_max, init=init, svd=svd,
                                           tol=tol, random_state=random_state)
    if return_errors:
        return (core, factors), errs
    return (core, factors)


# --------------------------------------------------------------------------
# Constrained CP by AO-ADMM  (decomposition/_constrained_cp.py,
# solvers/admm.py, tenalg/proximal.py)
# --------------------------------------------------------------------------
CONSTRAINT_NAMES = ("non_negative", "l1_reg", "l2_reg", "l2_square_reg", "unimodality", "normalize", "simplex",
                    "normalized_sparsity", "soft_sparsity", "smoothness", "monotonicity", "hard_sparsity")


def validate_constraints(n_const=1, order=0, **constraints):
    """``tenalg/proximal.py:4-150``: the constraint and parameter of mode
    ``order``.  Each keyword is None/False (unused), a scalar or True (every
    mode), a list (per mode, falsy entries unused) or a dict (mode ->
    parameter); a mode constrained twice raises ValueError.  Later names in
    the reference's order overwrite earlier ones for the modes they set."""
    unknown = set(constraints) - set(CONSTRAINT_NAMES)
    if unknown:
        raise TypeError(f"unknown constraints {sorted(unknown)}")
    taken = set()
    for name in CONSTRAINT_NAMES:
        spec = constraints.get(name)
        if not spec:
            continue
        if isinstance(spec, dict):
            modes = list(spec)
        elif isinstance(spec, list):
            modes = [m for m in range(len(spec)) if spec[m]]
        else:
            if taken:
                raise ValueError("You selected two constraints for the same mode. Consider to check your input")
            modes = list(range(n_const))
        for m in modes:
            if m in taken:
                raise ValueError("You selected two constraints for the same mode. Consider to check your input")
            taken.add(m)
    kinds = [None] * n_const
    params = [None] * n_const
    for name in CONSTRAINT_NAMES:
        spec = constraints.get(name)
        if not spec:
            continue
        if isinstance(spec, dict):
            for m, p in spec.items():
                kinds[m], params[m] = name, p
        else:
            for m in range(n_const):
                kinds[m] = name
                params[m] = spec[m] if isinstance(spec, list) else spec
    return kinds[order], params[order]


def soft_thresholding(tensor, threshold):
    """``tenalg/proximal.py:659-689``: sign(x) max(|x| - t, 0)."""
    return np.sign(tensor) * np.clip(np.abs(tensor) - threshold, 0, None)


def l2_square_prox(tensor, regularizer):
    """``tenalg/proximal.py:478-500``."""
    return tensor / (1 + 2 * regularizer)


def l2_prox(tensor, regularizer):
    """``tenalg/proximal.py:503-529``: block soft thresholding with the
    Frobenius norm of the whole matrix."""
    n = np.linalg.norm(tensor)
    big = n if n > regularizer else regularizer
    return tensor - tensor * regularizer / big


def hard_thresholding(tensor, number_of_non_zero):
    """``tenalg/proximal.py:634-656``: keep the ``number_of_non_zero``
    entries of largest magnitude (ranks from a double argsort of |x|, stable
    here so that equal magnitudes rank the later flat index first)."""
    v = np.ravel(tensor)
    order = np.argsort(np.abs(v), kind="stable")[::-1]
    ranks = np.empty(v.shape[0], dtype=np.int64)
    ranks[order] = np.arange(v.shape[0])
    return np.reshape(np.where(ranks < number_of_non_zero, v, 0.0), np.shape(tensor))


def normalized_sparsity_prox(tensor, threshold):
    """``tenalg/proximal.py:532-561``."""
    h = hard_thresholding(tensor, threshold)
    return h / np.linalg.norm(h)


def simplex_prox(tensor, parameter):
    """``tenalg/proximal.py:592-631``: each column projected on the simplex
    {x >= 0, sum x = parameter} by the sorted-cumsum rule (Held, Wolfe and
    Crowder): c_k = (sum of the k largest - parameter) / k, K = #{k: u_k >
    c_k}, x = max(v - c_K, 0)."""
    t = np.asarray(tensor, dtype=np.float64)
    one_d = t.ndim == 1
    if one_d:
        t = t[:, None]
    u = -np.sort(-t, axis=0)
    c = (np.cumsum(u, axis=0) - parameter) / np.arange(1, t.shape[0] + 1)[:, None]
    K = np.sum(u > c, axis=0) - 1
    shift = c[K, np.arange(t.shape[1])]
    out = np.clip(t - shift[None, :], 0, None)
    return out[:, 0] if one_d else out


def soft_sparsity_prox(tensor, threshold):
    """``tenalg/proximal.py:564-589``."""
    return simplex_prox(np.abs(tensor), threshold) * np.sign(tensor)


def smoothness_prox(tensor, regularizer):
    """``tenalg/proximal.py:270-291``: solve with the tridiagonal matrix
    (2 lambda + 1 on the diagonal, -lambda beside it)."""
    n = tensor.shape[0]
    A = (np.diag(np.full(n, 2 * regularizer + 1.0)) + np.diag(np.full(n - 1, -regularizer), 1)
         + np.diag(np.full(n - 1, -regularizer), -1))
    return np.linalg.solve(A, tensor)


def monotonicity_prox(tensor, decreasing=False):
    """``tenalg/proximal.py:294-356``, per column: y_j = the largest mean of
    x[i..j] over i <= j, floored at 0 for every j but the last (the
    reference's helper matrix holds zeros below its diagonal), then each y_j
    lowered to min(y_j, y_{j+1}) from the bottom up; ``decreasing`` works on
    the flipped column."""
    t = np.array(tensor, dtype=np.float64)
    if t.ndim == 1:
        t = t[:, None]
    elif t.ndim > 2:
        raise ValueError("Monotonicity prox doesn't support an input which has more than 2 dimensions.")
    if decreasing:
        t = t[::-1]
    n, ncol = t.shape
    cum = np.cumsum(t, axis=0)
    out = t.copy()
    for c in range(ncol):
        helper = np.zeros((n, n))
        for i in range(n):
            prev = cum[i - 1, c] if i > 0 else 0.0
            helper[i, i:] = (cum[i:, c] - prev) / (np.arange(n - i) + 1)
        col = np.max(helper, axis=0)
        for i in range(n - 2, -1, -1):
            if col[i] > col[i + 1]:
                col[i] = col[i + 1]
        out[:, c] = col
    if decreasing:
        out = out[::-1]
    return out


def unimodality_prox(tensor):
    """``tenalg/proximal.py:359-475``: per column, the peak j among entries
    at or above both monotone projections minimising sum_{k<j}|x_k - inc_k|
    + sum_{k>j}|x_k - dec_k| (columns with no such entry, or whose best cost
    is the global maximum, take j = 0); the result is inc above the peak,
    x at it and dec below it."""
    t = np.array(tensor, dtype=np.float64)
    if t.ndim == 1:
        t = t[:, None]
    elif t.ndim > 2:
        raise ValueError("Unimodality prox doesn't support an input which has more than 2 dimensions.")
    inc = monotonicity_prox(t)
    dec = monotonicity_prox(t, decreasing=True)
    peak = (t - dec >= 0) & (t - inc >= 0)
    a_inc = np.abs(t - inc)
    a_dec = np.abs(t - dec)
    before = np.where(peak, np.cumsum(a_inc, axis=0) - a_inc, 0.0)
    after_flipped = np.cumsum(a_dec[::-1], axis=0) - a_dec[::-1]
    after = np.where(peak, after_flipped[::-1], 0.0)
    cost = before + after
    cost = np.where(peak, cost, np.max(cost))
    jstar = np.argmin(cost, axis=0)
    out = t.copy()
    for c, j in enumerate(jstar):
        out[:j, c] = inc[:j, c]
        out[j + 1:, c] = dec[j + 1:, c]
    return out


def proximal_operator(tensor, n_const=1, order=0, **constraints):
    """``tenalg/proximal.py:153-267``."""
    if n_const is None:
        return tensor
    kind, p = validate_constraints(n_const=n_const, order=order, **constraints)
    if kind is None:
        return tensor
    if kind == "non_negative":
        # np.clip(x, 0, max(x)): when every entry is negative the bounds
        # cross and numpy returns the upper bound everywhere
        return np.minimum(np.maximum(tensor, 0), np.max(tensor))
    if kind == "l1_reg":
        return soft_thresholding(tensor, p)
    if kind == "l2_reg":
        return l2_prox(tensor, p)
    if kind == "l2_square_reg":
        return l2_square_prox(tensor, p)
    if kind == "unimodality":
        return unimodality_prox(tensor)
    if kind == "normalize":
        return tensor / np.max(np.abs(tensor))
    if kind == "simplex":
        return simplex_prox(tensor, p)
    if kind == "normalized_sparsity":
        return normalized_sparsity_prox(tensor, p)
    if kind == "soft_sparsity":
        return soft_sparsity_prox(tensor, p)
    if kind == "smoothness":
        return smoothness_prox(tensor, p)
    if kind == "monotonicity":
        return monotonicity_prox(tensor)
    if kind == "hard_sparsity":
        return hard_thresholding(tensor, p)
    raise RuntimeError("Invalid constraint name")


def admm(UtM, UtU, x, dual_var, n_iter_max=100, n_const=None, order=None, tol=1e-4, **constraints):
    """``solvers/admm.py:5-155``.  With rho = trace(UtU) / R:
    x_split = solve((UtU + rho I)^T, (UtM + rho (x + dual))^T),
    x = prox(x_split^T - dual), dual += x - x_split^T, stopping when
    ||x - x_split^T|| < tol ||x|| and ||x - x_old|| < tol ||dual||.
    Returns (x, x_split, dual)."""
    R = x.shape[1]
    rho = np.trace(UtU) / R
    x_split = None
    for _ in range(n_iter_max):
        x_old = x.copy()
        x_split = np.linalg.solve(np.transpose(UtU + rho * np.eye(UtU.shape[1])),
                                  np.transpose(UtM + rho * (x + dual_var)))
        x = proximal_operator(np.transpose(x_split) - dual_var, n_const=n_const, order=order, **constraints)
        if n_const is None:
            x = np.transpose(np.linalg.solve(np.transpose(UtU), np.transpose(UtM)))
            return x, x_split, dual_var
        dual_var = dual_var + x - np.transpose(x_split)
        dual_residual = x - np.transpose(x_split)
        primal_residual = x - x_old
        if (np.linalg.norm(dual_residual) < tol * np.linalg.norm(x)
                and np.linalg.norm(primal_residual) < tol * np.linalg.norm(dual_var)):
            break
    return x, x_split, dual_var


def initialize_constrained_parafac(tensor, rank, init="svd", svd="truncated_svd", random_state=None,
                                   **constraints):
    """``decomposition/_constrained_cp.py:22-163``.  ``init='random'`` draws
    from the global numpy RandomState (the reference passes no
    ``random_state`` to ``random_cp``, ``:104-107``); every factor then goes
    through its mode's proximal operator.  A CP tensor init returns its
    factors scaled by the weights' geometric mean, unprojected."""
    n_modes = tensor.ndim
    rng = check_random_state(random_state)
    if isinstance(init, str) and init == "random":
        _, factors = random_cp(tensor.shape, rank, normalise_factors=False, random_state=None,
                               dtype=tensor.dtype)
    elif isinstance(init, str) and init == "svd":
        factors = []
        for mode in range(n_modes):
            U, S, _ = svd_interface(unfold(tensor, mode), n_eigenvecs=rank, method=svd)
            if mode == 0:
                k = min(rank, S.shape[0])
                U = U.copy()
                U[:, :k] = U[:, :k] * S[:k]
            if tensor.shape[mode] < rank:
                pad = np.asarray(rng.random_sample((U.shape[0], rank - tensor.shape[mode])), dtype=tensor.dtype)
                U = np.concatenate([U, pad], axis=1)
            factors.append(U[:, :rank])
    elif isinstance(init, (tuple, list)):
        weights, factors = init
        factors = [np.array(f) for f in factors]
        if weights is None or np.all(weights == 1):
            return np.ones(rank, dtype=factors[0].dtype), factors
        avg = np.prod(weights) ** (1.0 / weights.shape[0])
        return np.ones(rank, dtype=factors[0].dtype), [f * avg for f in factors]
    else:
        raise ValueError(f'Initialization method "{init}" not recognized')
    factors = [proximal_operator(f, n_const=n_modes, order=i, **constraints) for i, f in enumerate(factors)]
    return np.ones(rank, dtype=tensor.dtype), factors


def constrained_parafac(tensor, rank, n_iter_max=100, n_iter_max_inner=10, init="svd", svd="truncated_svd",
                        tol_outer=1e-8, tol_inner=1e-6, random_state=None, verbose=0, return_errors=False,
                        cvg_criterion="abs_rec_error", fixed_modes=None, **constraints):
    """``decomposition/_constrained_cp.py:166-424``: per mode the MTTKRP
    with unit weights (``:352``) and one ADMM solve; the error from the last
    mode's MTTKRP and the Grams (``:379-383``); the outer loop stops on the
    summed constraint violation or on the error decrease (``:384-411``)."""
    rank = validate_cp_rank(tensor.shape, rank=rank)
    validate_constraints(n_const=tensor.ndim, **constraints)
    weights, factors = initialize_constrained_parafac(tensor, rank, init=init, svd=svd, random_state=random_state,
                                                      **constraints)
    factors = list(factors)
    rec_errors = []
    norm_tensor = norm(tensor, 2)
    fixed_modes = list(fixed_modes) if fixed_modes is not None else []
    if tensor.ndim - 1 in fixed_modes:
        warnings.warn("You asked for fixing the last mode, which is not supported.")
        fixed_modes.remove(tensor.ndim - 1)
    modes_list = [m for m in range(tensor.ndim) if m not in fixed_modes]
    duals = [np.zeros(f.shape) for f in factors]
    aux = [np.zeros(f.shape).T for f in factors]
    mttkrp = None
    for iteration in range(n_iter_max):
        for mode in modes_list:
            V = np.ones((rank, rank), dtype=tensor.dtype)
            for i, f in enumerate(factors):
                if i != mode:
                    V = V * np.dot(np.transpose(f), f)
            mttkrp = unfolding_dot_khatri_rao(tensor, (None, factors), mode)
            factors[mode], aux[mode], duals[mode] = admm(mttkrp, V, factors[mode], duals[mode],
                                                         n_iter_max=n_iter_max_inner, n_const=tensor.ndim,
                                                         order=mode, tol=tol_inner, **constraints)
        factors_norm = cp_norm((weights, factors))
        iprod = np.sum(np.sum(mttkrp * factors[-1], axis=0) * weights)
        rec_error = np.sqrt(np.abs(norm_tensor ** 2 + factors_norm ** 2 - 2 * iprod)) / norm_tensor
        rec_errors.append(rec_error)
        constraint_error = 0.0
        for mode in modes_list:
            constraint_error += np.linalg.norm(factors[mode] - np.transpose(aux[mode])) / np.linalg.norm(factors[mode])
        if tol_outer:
            if iteration >= 1:
                dec = rec_errors[-2] - rec_errors[-1]
                if constraint_error < tol_outer:
                    break
                if cvg_criterion == "abs_rec_error":
                    stop = abs(dec) < tol_outer
                elif cvg_criterion == "rec_error":
                    stop = dec < tol_outer
                else:
                    raise TypeError("Unknown convergence criterion")
                if stop:
                    break
    if return_errors:
        return (weights, factors), rec_errors
    return (weights, factors)
