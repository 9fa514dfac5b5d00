"""Non-negative CP by multiplicative updates (``tensorly/decomposition/_nn_cp.py``,
0.9.0) on the gfx950 kernels.

:func:`non_negative_parafac` keeps the reference signature (``_nn_cp.py:26-40``)
and return convention (``:174-180``).  Per mode (``:111-135``) it runs the
fused MTTKRP (``tla_mttkrp``) and one ``tla_nn_mu_update`` launch, which forms
``accum`` = (Hadamard of the other modes' FP64 Grams) * w w^T and applies
``A <- A * max(M, eps) / max(A accum, eps)`` in place; the updated factor's
FP64 Gram comes from ``tla_gram64``.  The error (``:138-150``) uses the last
mode's MTTKRP as in CP-ALS.  ``init='svd'`` is the NNDSVDa initialisation of
the reference (``tenalg/svd.py:68-127``) on the GPU truncated SVD.

Not implemented (raise NotImplementedError rather than run elsewhere):
``mask``.
"""
import warnings

import numpy as np
import torch

from .. import _lib
from .._lib import check, lib, ptr, ptr_array, stream
from ..cp_tensor import CPTensor, cp_normalize, grams64, validate_cp_rank
from ..tenalg import unfolding_dot_khatri_rao
from ._cp import error_calc, initialize_cp
from ._tucker import _sqnorm64

__all__ = ["non_negative_parafac", "non_negative_parafac_hals"]


def _mu_update(factors, mttkrp, grams, weights, mode, eps):
    f = factors[mode]
    rows, rank = f.shape
    m = mttkrp.contiguous()
    gp, _keep = ptr_array(grams)
    check(lib().tla_nn_mu_update(_lib.dtype_code(f), ptr(f), ptr(m), rows, rank, gp, len(factors), mode,
                                 ptr(weights), _lib.dtype_code(weights), float(eps), stream(f.device)),
          "nn_mu_update")


def non_negative_parafac(tensor, rank, n_iter_max=100, init="svd", svd="truncated_svd", tol=10e-7,
                         random_state=None, verbose=0, normalize_factors=False, return_errors=False, mask=None,
                         cvg_criterion="abs_rec_error", fixed_modes=None):
    """Non-negative CP decomposition by multiplicative updates
    (``_nn_cp.py:26-180``); ``tensor`` is a float32 / float64 GPU tensor."""
    _lib.require_gpu_tensor(tensor, "tensor")
    if mask is not None:
        raise NotImplementedError("mask is not supported by tensorly_amd.non_negative_parafac")
    if tensor.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"non_negative_parafac takes float32 / float64 tensors, got {tensor.dtype}")
    tensor = tensor.contiguous()
    epsilon = torch.finfo(tensor.dtype).eps
    rank = validate_cp_rank(tuple(tensor.shape), rank=rank)
    weights, factors = initialize_cp(tensor, rank, init=init, svd=svd, non_negative=True, random_state=random_state,
                                     normalize_factors=normalize_factors)
    if weights is None:
        weights = torch.ones(rank, dtype=tensor.dtype, device=tensor.device)
    weights = weights.to(tensor.dtype).contiguous()
    factors = [f.to(tensor.dtype).contiguous().clone() for f in factors]
    rec_errors = []
    norm_tensor = torch.sqrt(_sqnorm64(tensor))

    fixed_modes = [] if fixed_modes is None else list(fixed_modes)
    if tensor.dim() - 1 in fixed_modes:
        warnings.warn("You asked for fixing the last mode, which is not supported while tol is fixed.\n "
                      "The last mode will not be fixed. Consider using tl.moveaxis()")
        fixed_modes.remove(tensor.dim() - 1)
    modes_list = [mode for mode in range(tensor.dim()) if mode not in fixed_modes]

    grams = grams64(factors)
    mttkrp = None
    for iteration in range(n_iter_max):
        if verbose > 1:
            print("Starting iteration", iteration + 1)
        for mode in modes_list:
            if verbose > 1:
                print("Mode", mode, "of", tensor.dim())
            mttkrp = unfolding_dot_khatri_rao(tensor, (weights, factors), mode)
            _mu_update(factors, mttkrp, grams, weights, mode, epsilon)
            if normalize_factors and mode != modes_list[-1]:
                weights, factors = cp_normalize((weights, factors))
                weights = weights.contiguous()
                factors = [f.contiguous() for f in factors]
                grams = grams64(factors)
            else:
                grams[mode] = grams64([factors[mode]])[0]

        if tol:
            unnorml_rec_error, _, _ = error_calc(tensor, norm_tensor, weights, factors, mttkrp=mttkrp)
            rec_error = float(unnorml_rec_error / norm_tensor)
            rec_errors.append(rec_error)
            if iteration >= 1:
                rec_error_decrease = rec_errors[-2] - rec_errors[-1]
                if verbose:
                    print(f"iteration {iteration}, reconstraction error: {rec_error}, "
                          f"decrease = {rec_error_decrease}")
                if cvg_criterion == "abs_rec_error":
                    stop_flag = abs(rec_error_decrease) < tol
                elif cvg_criterion == "rec_error":
                    stop_flag = rec_error_decrease < tol
                else:
                    raise TypeError("Unknown convergence criterion")
                if stop_flag:
                    if verbose:
                        print(f"PARAFAC converged after {iteration} iterations")
                    break
            elif verbose:
                print(f"reconstruction error={rec_errors[-1]}")
        if normalize_factors:
            weights, factors = cp_normalize((weights, factors))
            weights = weights.contiguous()
            factors = [f.contiguous() for f in factors]
            grams = grams64(factors)
    cp_tensor = CPTensor((weights, factors))
    if return_errors:
        return cp_tensor, rec_errors
    return cp_tensor


def _hals_update(factors, mttkrp, grams, weights, mode, sparsity, n_iter_max, tol):
    f = factors[mode]
    rows, rank = f.shape
    m = mttkrp.contiguous()
    gp, _keep = ptr_array(grams)
    nb = lib().tla_hals_nnls_workspace_size(rows, rank)
    ws = _lib.workspace(nb, f.device)
    check(lib().tla_hals_nnls(_lib.dtype_code(f), ptr(f), ptr(m), rows, rank, gp, len(factors), mode, ptr(weights),
                              _lib.dtype_code(weights), 0.0 if sparsity is None else float(sparsity),
                              0 if sparsity is None else 1, int(n_iter_max), float(tol), None, ptr(ws), nb,
                              stream(f.device)),
          "hals_nnls")


def _solve_update(factors, mttkrp, grams, weights, mode, seen):
    """Unconstrained mode of the HALS driver: A = M (pseudo_inverse)^-1, the
    FP64 ALS solve (``tla_cp_solve``)."""
    f = factors[mode]
    rows, rank = f.shape
    m = mttkrp.contiguous()
    gp, _keep = ptr_array(grams)
    g_out = torch.empty((rank, rank), dtype=torch.float64, device=f.device)
    status = torch.zeros(1, dtype=torch.int32, device=f.device)
    nb = lib().tla_cp_solve_workspace_size(rows, rank)
    ws = _lib.workspace(nb, f.device)
    check(lib().tla_cp_solve(_lib.dtype_code(f), gp, len(factors), mode, ptr(weights), 0.0, ptr(m), rows, rank,
                             ptr(f), ptr(g_out), None, ptr(status), None, ptr(ws), nb, stream(f.device)),
          "cp_solve")
    torch.maximum(seen, status, out=seen)
    return g_out


def _raise_if_singular(seen):
    """``tl.solve`` raises on a singular system (``_nn_cp.py:338``)."""
    if int(seen) == 2:
        raise np.linalg.LinAlgError("Singular matrix")


def non_negative_parafac_hals(tensor, rank, n_iter_max=100, init="svd", svd="truncated_svd", tol=10e-8,
                              random_state=None, sparsity_coefficients=None, fixed_modes=None, nn_modes="all",
                              exact=False, normalize_factors=False, verbose=False, return_errors=False,
                              cvg_criterion="abs_rec_error"):
    """Non-negative CP decomposition by HALS (``_nn_cp.py:183-360``).  Per
    mode: the fused MTTKRP, then one ``tla_hals_nnls`` launch running the
    whole ``hals_nnls`` loop of ``solvers/nnls.py:5-173`` (``n_iter_max=100``,
    or 50000 iterations and ``tol=1e-16`` when ``exact``) for non-negative
    modes, or the FP64 ALS solve for the others."""
    _lib.require_gpu_tensor(tensor, "tensor")
    if tensor.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"non_negative_parafac_hals takes float32 / float64 tensors, got {tensor.dtype}")
    tensor = tensor.contiguous()
    weights, factors = initialize_cp(tensor, rank, init=init, svd=svd, non_negative=True, random_state=random_state,
                                     normalize_factors=normalize_factors)
    if weights is None:
        weights = torch.ones(rank, dtype=tensor.dtype, device=tensor.device)
    weights = weights.to(tensor.dtype).contiguous()
    factors = [f.to(tensor.dtype).contiguous().clone() for f in factors]
    norm_tensor = torch.sqrt(_sqnorm64(tensor))
    n_modes = tensor.dim()
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
    for mode in range(n_modes):
        if sparsity_coefficients[mode] is not None:
            warnings.warn("Sparsity coefficient is ignored in unconstrained modes.")
    modes = [mode for mode in range(n_modes) if mode not in fixed_modes]
    inner_iters, inner_tol = (50000, 1e-16) if exact else (100, 1e-8)

    rec_errors = []
    grams = grams64(factors)
    mttkrp = None
    seen = torch.zeros(1, dtype=torch.int32, device=tensor.device)
    for iteration in range(n_iter_max):
        for mode in modes:
            mttkrp = unfolding_dot_khatri_rao(tensor, (weights, factors), mode)
            if mode in nn_modes:
                _hals_update(factors, mttkrp, grams, weights, mode, sparsity_coefficients[mode], inner_iters,
                             inner_tol)
                grams[mode] = grams64([factors[mode]])[0]
            else:
                grams[mode] = _solve_update(factors, mttkrp, grams, weights, mode, seen)
            if normalize_factors and mode != modes[-1]:
                weights, factors = cp_normalize((weights, factors))
                weights = weights.contiguous()
                factors = [f.contiguous() for f in factors]
                grams = grams64(factors)
        if tol:
            _raise_if_singular(seen)
            unnorml_rec_error, _, _ = error_calc(tensor, norm_tensor, weights, factors, mttkrp=mttkrp)
            rec_error = float(unnorml_rec_error / norm_tensor)
            rec_errors.append(rec_error)
            if iteration >= 1:
                rec_error_decrease = rec_errors[-2] - rec_errors[-1]
                if verbose:
                    print(f"iteration {iteration}, reconstruction error: {rec_error}, "
                          f"decrease = {rec_error_decrease}")
                if cvg_criterion == "abs_rec_error":
                    stop_flag = abs(rec_error_decrease) < tol
                elif cvg_criterion == "rec_error":
                    stop_flag = rec_error_decrease < tol
                else:
                    raise TypeError("Unknown convergence criterion")
                if stop_flag:
                    if verbose:
                        print(f"PARAFAC converged after {iteration} iterations")
                    break
            elif verbose:
                print(f"reconstruction error={rec_errors[-1]}")
        if normalize_factors:
            weights, factors = cp_normalize((weights, factors))
            weights = weights.contiguous()
            factors = [f.contiguous() for f in factors]
            grams = grams64(factors)
    _raise_if_singular(seen)
    cp_tensor = CPTensor((weights, factors))
    if return_errors:
        return cp_tensor, rec_errors
    return cp_tensor
