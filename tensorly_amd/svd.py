"""SVD interface on the GPU (``tensorly/tenalg/svd.py``, 0.9.0).

The leading left singular vectors of ``A`` (m x q) are taken as the leading
eigenvectors of the FP64 Gram ``A A^T`` (or of ``unfold(X, n) unfold(X, n)^T``
for a tensor unfolding, formed without the copy), computed by the
``tla_sym_eig`` kernel; singular values are the square roots of the
eigenvalues.  :func:`svd_flip` applies the reference sign rule
(``svd.py:13-65``): the largest-magnitude entry of each U column (first one
on ties) is made positive.
"""
import warnings

import torch

from . import _lib
from ._lib import check, lib, ptr, shape_array, stream, workspace

__all__ = ["svd_flip", "svd_checks", "truncated_svd", "svd_interface", "make_svd_non_negative", "gram_eigh",
           "gram_left_singular", "qr",
           "mode_gram", "SVD_FUNS"]

SVD_FUNS = ["truncated_svd", "symeig_svd"]

def mode_gram(tensor, mode):
    """FP64 ``unfold(X, mode) @ unfold(X, mode)^T`` without forming the
    unfolding."""
    _lib.require_gpu_tensor(tensor, "tensor")
    x = tensor.contiguous()
    n = x.shape[mode]
    sh, _k = shape_array(x.shape)
    nb = lib().tla_mode_gram64_workspace_size(x.dim(), sh, mode)
    ws = workspace(nb, x.device)
    G = torch.empty((n, n), dtype=torch.float64, device=x.device)
    check(lib().tla_mode_gram64(_lib.dtype_code(x), ptr(x), x.dim(), sh, mode, ptr(G), ptr(ws), nb,
                                stream(x.device)), "mode_gram64")
    return G


def gram_eigh(G, k, warm=None, tol=1e-11, max_iter=300):
    """Leading ``k`` eigenpairs (descending) of the symmetric FP64 ``G``:
    returns ``(evals (k,), evecs (n, k))`` as FP64 GPU tensors."""
    n = G.shape[0]
    G = G.to(torch.float64).contiguous()
    evals = torch.empty(k, dtype=torch.float64, device=G.device)
    evecs = torch.empty((n, k), dtype=torch.float64, device=G.device)
    q0 = warm.to(torch.float64).contiguous() if warm is not None else None
    nb = lib().tla_sym_eig_workspace_size(n)
    ws = workspace(nb, G.device)
    check(lib().tla_sym_eig(ptr(G), n, k, ptr(q0), q0.shape[1] if q0 is not None else 0, ptr(evals), ptr(evecs),
                            float(tol), int(max_iter), None, ptr(ws), nb, stream(G.device)), "sym_eig")
    return evals, evecs


def gram_eigh_batch(Gs, ks, warms=None, tol=1e-11, max_iter=300):
    """:func:`gram_eigh` of independent Grams side by side: one launch per
    solver kind, one workgroup per Gram (``tla_sym_eig_batch``).  Returns a
    list of ``(evals, evecs)``."""
    out = []
    for c0 in range(0, len(Gs), 8):
        chunk = list(range(c0, min(len(Gs), c0 + 8)))
        dev = Gs[chunk[0]].device
        G = [Gs[i].to(torch.float64).contiguous() for i in chunk]
        n = [int(g.shape[0]) for g in G]
        k = [int(ks[i]) for i in chunk]
        q0 = [warms[i].to(torch.float64).contiguous() if warms is not None and warms[i] is not None else None
              for i in chunk]
        evals = [torch.empty(kk, dtype=torch.float64, device=dev) for kk in k]
        evecs = [torch.empty((nn, kk), dtype=torch.float64, device=dev) for nn, kk in zip(n, k)]
        nbs = [int(lib().tla_sym_eig_workspace_size(nn)) for nn in n]
        wss = [workspace(nb, dev) for nb in nbs]
        Ip = _lib.ctypes.c_int * len(chunk)
        gp, _kg = _lib.ptr_array(G)
        qp, _kq = _lib.ptr_array(q0)
        ep, _ke = _lib.ptr_array(evals)
        vp, _kv = _lib.ptr_array(evecs)
        wp, _kw = _lib.ptr_array(wss)
        nbarr = (_lib.ctypes.c_size_t * len(chunk))(*nbs)
        k0 = Ip(*[q.shape[1] if q is not None else 0 for q in q0])
        check(lib().tla_sym_eig_batch(len(chunk), gp, Ip(*n), Ip(*k), qp, k0, ep, vp, float(tol), int(max_iter),
                                      wp, nbarr, stream(dev)), "sym_eig_batch")
        out.extend(zip(evals, evecs))
    return out


def svd_flip(U, V, u_based_decision=True):
    """Sign rule of ``svd.py:13-65`` (on GPU tensors)."""
    if u_based_decision:
        idx = torch.argmax(torch.abs(U), dim=0)
        signs = torch.sign(U[idx, torch.arange(U.shape[1], device=U.device)])
        U = U * signs
        if V is not None:
            if V.shape[0] > U.shape[1]:
                signs = torch.cat((signs, torch.ones(V.shape[0] - U.shape[1], dtype=V.dtype, device=V.device)))
            V = V * signs[: V.shape[0]][:, None]
    else:
        idx = torch.argmax(torch.abs(V), dim=1)
        signs = torch.sign(V[torch.arange(V.shape[0], device=V.device), idx])
        V = V * signs[:, None]
        if U.shape[1] > V.shape[0]:
            signs = torch.cat((signs, torch.ones(U.shape[1] - V.shape[0], dtype=V.dtype, device=V.device)))
        U = U * signs[: U.shape[1]]
    return U, V


def svd_checks(matrix, n_eigenvecs=None):
    """``svd.py:173-208``."""
    if matrix.dim() != 2:
        raise ValueError(f"matrix be a matrix. matrix.ndim is {matrix.dim()} != 2")
    d1, d2 = matrix.shape
    lo, hi = min(d1, d2), max(d1, d2)
    if n_eigenvecs is None:
        n_eigenvecs = hi
    if n_eigenvecs > hi:
        warnings.warn(f"Trying to compute SVD with n_eigenvecs={n_eigenvecs}, which is larger "
                      f"than max(matrix.shape)={hi}. Setting n_eigenvecs to {hi}.")
        n_eigenvecs = hi
    return n_eigenvecs, lo, hi


def _left_svd_from_gram(G, n_rows, min_dim, n_eig, warm=None):
    """U (n_rows x min(n_eig, n_rows)), S (min(n_eig, min_dim)) from the row
    Gram."""
    k = min(n_eig, n_rows)
    evals, evecs = gram_eigh(G, k, warm=warm)
    s = torch.sqrt(torch.clamp(evals, min=0.0))
    return evecs, s[: min(n_eig, min_dim)]


def truncated_svd(matrix, n_eigenvecs=None, warm=None, **kwargs):
    """``svd.py:211-235`` semantics: U (m x n_eig), S, V (n_eig x q) with
    ``V = diag(1/S) U^T A`` (rows with S == 0 left zero)."""
    n_eig, lo, _ = svd_checks(matrix, n_eigenvecs)
    A = matrix.contiguous()
    G = mode_gram(A, 0)
    U, S = _left_svd_from_gram(G, A.shape[0], lo, n_eig, warm=warm)
    U = U.to(A.dtype)
    k = S.shape[0]
    from .tenalg import mode_dot
    V = mode_dot(A, U[:, :k], 0, transpose=True)  # k x q
    inv = torch.where(S > 0, 1.0 / torch.where(S > 0, S, torch.ones_like(S)), torch.zeros_like(S))
    V = V * inv.to(V.dtype)[:, None]
    return U, S.to(A.dtype), V


def make_svd_non_negative(tensor, U, S, V, nntype=True):
    """NNDSVD / NNDSVDa of ``svd.py:68-127`` on the GPU tensors of a
    truncated SVD (column loop over the rank; init-time only)."""
    if nntype is True:
        nntype = "nndsvda"
    if nntype not in ("nndsvd", "nndsvda"):
        raise ValueError(f'Invalid nntype parameter: got {nntype} instead of one of ("nndsvd", "nndsvda")')
    W = torch.zeros_like(U)
    H = torch.zeros_like(V)
    W[:, 0] = torch.sqrt(S[0]) * torch.abs(U[:, 0])
    H[0, :] = torch.sqrt(S[0]) * torch.abs(V[0, :])
    for j in range(1, min(U.shape[1], V.shape[0])):
        x, y = U[:, j], V[j, :]
        x_p, y_p = torch.clamp(x, min=0.0), torch.clamp(y, min=0.0)
        x_n, y_n = torch.abs(torch.clamp(x, max=0.0)), torch.abs(torch.clamp(y, max=0.0))
        x_p_nrm, y_p_nrm = torch.linalg.vector_norm(x_p), torch.linalg.vector_norm(y_p)
        x_n_nrm, y_n_nrm = torch.linalg.vector_norm(x_n), torch.linalg.vector_norm(y_n)
        m_p, m_n = x_p_nrm * y_p_nrm, x_n_nrm * y_n_nrm
        if bool(m_p > m_n):
            u, v, sigma = x_p / x_p_nrm, y_p / y_p_nrm, m_p
        else:
            u, v, sigma = x_n / x_n_nrm, y_n / y_n_nrm, m_n
        lbd = torch.sqrt(S[j] * sigma)
        W[:, j] = lbd * u
        H[j, :] = lbd * v
    eps = torch.finfo(tensor.dtype).eps
    if nntype == "nndsvd":
        W = torch.sign(W) * torch.clamp(torch.abs(W) - eps, min=0.0)
        H = torch.sign(H) * torch.clamp(torch.abs(H) - eps, min=0.0)
    else:
        avg = torch.mean(tensor)
        W = torch.where(W < eps, avg.to(W.dtype), W)
        H = torch.where(H < eps, avg.to(H.dtype), H)
    return W, H


def svd_interface(matrix, method="truncated_svd", n_eigenvecs=None, flip_sign=True, u_based_flip_sign=True,
                  non_negative=None, mask=None, n_iter_mask_imputation=5, **kwargs):
    """``svd.py:366-447``.  With ``mask`` (and ``n_eigenvecs``) the missing
    entries are filled ``n_iter_mask_imputation`` times from the current
    truncated SVD, U diag(S) V (two MFMA products), by ``tla_masked_blend``,
    and the SVD recomputed (``:431-439``); ``non_negative`` applies
    :func:`make_svd_non_negative` after the sign flip, as the reference."""
    if method not in ("truncated_svd", "symeig_svd") and not callable(method):
        raise ValueError(f"Got svd={method}. However, the possible choices are {SVD_FUNS} or to pass a callable.")
    fun = method if callable(method) else truncated_svd
    U, S, V = fun(matrix, n_eigenvecs=n_eigenvecs, **kwargs)
    if mask is not None and n_eigenvecs is not None:
        from .tenalg import mode_dot
        _lib.require_on_device(mask, matrix.device, "mask")
        m = mask.to(matrix.dtype).contiguous()
        for _ in range(n_iter_mask_imputation):
            k = S.shape[0]
            low = mode_dot((V[:k] * S[:, None]).contiguous(), U[:, :k].contiguous(), 0)
            blended = torch.empty_like(low)
            x = matrix.contiguous()
            check(lib().tla_masked_blend(_lib.dtype_code(x), ptr(x), ptr(m), ptr(low), ptr(blended), x.numel(),
                                         stream(x.device)), "masked_blend")
            matrix = blended
            U, S, V = fun(matrix, n_eigenvecs=n_eigenvecs, **kwargs)
    if flip_sign:
        U, V = svd_flip(U, V, u_based_decision=u_based_flip_sign)
    if non_negative is not False and non_negative is not None:
        U, V = make_svd_non_negative(matrix, U, S, V, non_negative)
    return U, S, V


def gram_left_singular(G, k, dtype, warm=None, want_block=False, tol=1e-11, max_iter=300):
    """Leading ``k`` left singular vectors (``svd_flip`` signs, ``dtype``) and
    singular values of the matrix whose FP64 row Gram is ``G``, in one
    ``tla_gram_left_singular`` launch.  ``warm``: FP64 warm-start columns,
    or the block an earlier call returned.  Returns ``(U, S, block)``;
    ``block`` (``want_block``, sizes with a subspace block only) is that
    call's final FP64 block, the warm start for a nearby matrix."""
    n = int(G.shape[0])
    G = G.to(torch.float64).contiguous()
    b = int(lib().tla_gram_svd_block(n, k))
    U = torch.empty((n, k), dtype=dtype, device=G.device)
    S = torch.empty(k, dtype=dtype, device=G.device)
    q0 = warm.to(torch.float64).contiguous() if (warm is not None and b > 0) else None
    blk = None
    if want_block and b > 0:
        blk = q0 if (q0 is not None and tuple(q0.shape) == (n, b)) else torch.empty((n, b), dtype=torch.float64,
                                                                                    device=G.device)
    nb = lib().tla_sym_eig_workspace_size(n)
    ws = workspace(nb, G.device)
    check(lib().tla_gram_left_singular(ptr(G), n, k, ptr(q0), q0.shape[1] if q0 is not None else 0,
                                       _lib.dtype_code(U), ptr(U), ptr(S), None, ptr(blk), float(tol), int(max_iter),
                                       None, ptr(ws), nb, stream(G.device)), "gram_left_singular")
    return U, S, blk


def leading_left_singular_vectors(tensor, mode, k, warm=None, return_warm=False):
    """Leading ``k`` left singular vectors of ``unfold(tensor, mode)`` with the
    reference sign rule, and the singular values -- without forming the
    unfolding (the HOOI path).  ``return_warm``: also the solver's final
    FP64 block (None for modes of 128 or fewer rows), the warm start of the
    next call on the same mode."""
    n = tensor.shape[mode]
    q = tensor.numel() // max(n, 1)
    G = mode_gram(tensor, mode)
    U, S, blk = gram_left_singular(G, min(k, n), tensor.dtype, warm=warm, want_block=return_warm)
    S = S[: min(k, n, q)]
    if return_warm:
        return U, S, blk
    return U, S


def leading_left_singular_vectors_batch(tensor, modes, ks):
    """:func:`leading_left_singular_vectors` of several modes of one tensor,
    cold-started, the eigensolves in one batched launch (the Tucker SVD
    init): a list of ``(U, S, U64)``, U64 the FP64 vectors (the first warm
    start of each mode's HOOI solves)."""
    Gs, kk, dims = [], [], []
    for mode, k in zip(modes, ks):
        n = tensor.shape[mode]
        q = tensor.numel() // max(n, 1)
        Gs.append(mode_gram(tensor, mode))
        kk.append(min(k, n))
        dims.append(min(n, q))
    res = []
    for (evals, evecs), k, md in zip(gram_eigh_batch(Gs, kk), ks, dims):
        s = torch.sqrt(torch.clamp(evals, min=0.0))[: min(k, md)]
        U, _ = svd_flip(evecs, None)
        res.append((U.to(tensor.dtype), s.to(tensor.dtype), U))
    return res


def qr(matrix):
    """Reduced QR ``(Q, R)`` of a tall GPU matrix (rows >= cols): ``tl.qr``
    of the numpy backend (``backend/numpy_backend.py:75``, ``np.linalg.qr``,
    LAPACK geqrf + orgqr) with the same Householder signs, in one
    ``tla_qr`` launch (FP64 arithmetic)."""
    _lib.require_gpu_tensor(matrix, "matrix")
    if matrix.dim() != 2:
        raise ValueError(f"qr takes a matrix, got {matrix.dim()} dimensions")
    if matrix.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"qr takes float32 / float64 matrices, got {matrix.dtype}")
    A = matrix.contiguous()
    rows, cols = A.shape
    if rows < cols:
        raise NotImplementedError("qr of a wide matrix (rows < cols) is not supported")
    Q = torch.empty_like(A)
    R = torch.empty((cols, cols), dtype=A.dtype, device=A.device)
    nb = lib().tla_qr_workspace_size(rows, cols)
    ws = workspace(nb, A.device)
    check(lib().tla_qr(_lib.dtype_code(A), ptr(A), rows, cols, ptr(Q), ptr(R), ptr(ws), nb, stream(A.device)), "qr")
    return Q, R
