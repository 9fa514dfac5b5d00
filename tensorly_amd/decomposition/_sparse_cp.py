"""CP-ALS on a sparse tensor in coordinate (COO) form: the role of
``tensorly/contrib/sparse/decomposition.py:16`` (``parafac`` of
``decomposition/_cp.py:230-543`` run over the pydata ``sparse`` backend),
without densifying.

Each mode update is the COO MTTKRP (``tla_mttkrp_coo``) followed by the FP64
Hadamard-of-Grams solve (``tla_cp_solve``, which also returns the new
factor's Gram); the error is ``sqrt(|‖X‖² + ‖M‖² − 2<X, M>|) / ‖X‖`` from the
last mode's MTTKRP (``_cp.py:217-225``), with ‖X‖² from the coalesced values
(repeated coordinates add up).  ``init='random'`` draws the factors as the
dense path does (``random/base.py:103-105``), a CP tensor is taken as the
dense ``initialize_cp`` takes it; ``init='svd'``, ``mask``,
``sparsity``, ``linesearch`` and ``normalize_factors`` raise
NotImplementedError.
"""
import numpy as np
import torch

from .. import _lib
from .._lib import check, lib, ptr, ptr_array, stream
from ..cp_tensor import CPTensor, grams64, validate_cp_rank
from ..tenalg import unfolding_dot_khatri_rao_coo
from ._cp import _check_random_state, initialize_cp

__all__ = ["parafac_coo"]


def _solve(factors, mttkrp, grams, weights, mode, status):
    """One mode update; ``status`` (device int32) receives 0 / 1 / 2 (2:
    singular normal equations)."""
    f = factors[mode]
    rows, rank = f.shape
    gp, _keep = ptr_array(grams)
    g_out = torch.empty((rank, rank), dtype=torch.float64, device=f.device)
    nb = lib().tla_cp_solve_workspace_size(rows, rank)
    ws = _lib.workspace(nb, f.device)
    check(lib().tla_cp_solve(_lib.dtype_code(f), gp, len(factors), mode, ptr(weights), 0.0, ptr(mttkrp.contiguous()),
                             rows, rank, ptr(f), ptr(g_out), None, ptr(status), None, ptr(ws), nb, stream(f.device)),
          "cp_solve")
    return g_out


def _raise_if_singular(status):
    """``tl.solve`` raises on a singular system (``_cp.py:425-427``)."""
    if int(torch.max(status)) == 2:
        raise np.linalg.LinAlgError("Singular matrix")


def parafac_coo(coords, values, shape, rank, n_iter_max=100, init="random", tol=1e-8, random_state=None,
                return_errors=False, cvg_criterion="abs_rec_error"):
    """CP-ALS of the sparse tensor ``(coords (ndim x nnz), values (nnz),
    shape)``; returns ``CPTensor`` (and the per-iteration relative errors
    when ``return_errors``), as ``parafac``."""
    _lib.require_gpu_tensor(values, "values")
    if values.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"parafac_coo takes float32 / float64 values, got {values.dtype}")
    if isinstance(init, str) and init != "random":
        raise NotImplementedError("parafac_coo supports init='random' or a CP tensor (an SVD init would densify)")
    shape = tuple(int(s) for s in shape)
    rank = validate_cp_rank(shape, rank=rank)
    dt, dev = values.dtype, values.device
    coords = coords.to(device=dev, dtype=torch.int64).contiguous()
    values = values.contiguous()
    if isinstance(init, str):
        rng = _check_random_state(random_state)
        factors = [torch.tensor(rng.random_sample((s, rank)), dtype=dt, device=dev) for s in shape]
    else:
        # a CP tensor: the weights are folded into the factors as the dense
        # initialize_cp does (_cp.py:102-121)
        _, factors = initialize_cp(values, rank, init=init)
        factors = [f.to(device=dev, dtype=dt).contiguous().clone() for f in factors]
    weights = torch.ones(rank, dtype=dt, device=dev)
    coalesced = torch.sparse_coo_tensor(coords, values, shape).coalesce().values()
    norm_tensor = torch.sqrt(torch.sum(coalesced.to(torch.float64) ** 2))
    grams = grams64(factors)
    rec_errors = []
    status = torch.zeros(len(shape), dtype=torch.int32, device=dev)
    seen = torch.zeros(len(shape), dtype=torch.int32, device=dev)
    for iteration in range(n_iter_max):
        mttkrp = None
        for mode in range(len(shape)):
            mttkrp = unfolding_dot_khatri_rao_coo(coords, values, shape, (weights, factors), mode)
            grams[mode] = _solve(factors, mttkrp, grams, weights, mode, status[mode:mode + 1])
        torch.maximum(seen, status, out=seen)
        if tol or return_errors:
            _raise_if_singular(seen)
            acc = grams[0]
            for g in grams[1:]:
                acc = acc * g
            nm2 = torch.sum(acc)  # weights are ones
            ip = torch.sum(mttkrp.to(torch.float64) * factors[-1].to(torch.float64))
            err = torch.sqrt(torch.abs(norm_tensor ** 2 + nm2 - 2 * ip)) / norm_tensor
            rec_errors.append(float(err))
            if tol and iteration >= 1:
                dec = rec_errors[-2] - rec_errors[-1]
                if cvg_criterion == "abs_rec_error":
                    stop = abs(dec) < tol
                elif cvg_criterion == "rec_error":
                    stop = dec < tol
                else:
                    raise TypeError("Unknown convergence criterion")
                if stop:
                    break
    _raise_if_singular(seen)
    cp = CPTensor((weights, factors))
    if return_errors:
        return cp, rec_errors
    return cp
