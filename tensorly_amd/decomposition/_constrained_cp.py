"""Constrained CP by AO-ADMM (``tensorly/decomposition/_constrained_cp.py``,
0.9.0) on the gfx950 kernels.

:func:`constrained_parafac` keeps the reference signature (``:166-189``) and
return convention (``:418-423``).  Per mode (``:341-375``) it runs the fused
MTTKRP with unit weights (``tla_mttkrp``) and one ``tla_admm_update`` launch
that forms UtU from the FP64 Grams and runs the whole inner ADMM loop with
the mode's proximal operator; the updated factor's Gram comes from
``tla_gram64``.  The error (``:377-383``) is formed on the device from the
Grams and the last mode's ``sum(UtM * x)``; the host reads errors only when
``tol_outer`` needs them each iteration (or once at the end for
``return_errors``).  A singular UtU + rho I raises
``numpy.linalg.LinAlgError`` as the reference's ``tl.solve`` does.
"""
import warnings

import numpy as np
import torch

from .. import _lib
from .._lib import check, lib, ptr, ptr_array, stream
from ..cp_tensor import CPTensor, grams64, validate_cp_rank
from ..proximal import _code_and_param, admm_update, proximal_operator, validate_constraints
from ..tenalg import unfolding_dot_khatri_rao
from ._cp import _check_random_state
from ._tucker import _sqnorm64

__all__ = ["constrained_parafac", "initialize_constrained_parafac", "ConstrainedCP"]

_CON_KW = ("non_negative", "l1_reg", "l2_reg", "l2_square_reg", "unimodality", "normalize", "simplex",
           "normalized_sparsity", "soft_sparsity", "smoothness", "monotonicity", "hard_sparsity")


def initialize_constrained_parafac(tensor, rank, init="svd", svd="truncated_svd", random_state=None,
                                   non_negative=None, l1_reg=None, l2_reg=None, l2_square_reg=None, unimodality=None,
                                   normalize=None, simplex=None, normalized_sparsity=None, soft_sparsity=None,
                                   smoothness=None, monotonicity=None, hard_sparsity=None):
    """``_constrained_cp.py:22-163``.  ``init='random'`` draws
    ``random_sample((I_m, R))`` per mode from numpy's global RandomState (the
    reference passes no ``random_state`` to ``random_cp``, ``:104-107``);
    ``init='svd'`` uses the GPU truncated SVD; every factor then goes through
    its mode's proximal operator.  A CP tensor init returns its factors
    scaled by the weights' geometric mean, unprojected."""
    cons = dict(non_negative=non_negative, l1_reg=l1_reg, l2_reg=l2_reg, l2_square_reg=l2_square_reg,
                unimodality=unimodality, normalize=normalize, simplex=simplex,
                normalized_sparsity=normalized_sparsity, soft_sparsity=soft_sparsity, smoothness=smoothness,
                monotonicity=monotonicity, hard_sparsity=hard_sparsity)
    n_modes = tensor.dim()
    dt, dev = tensor.dtype, tensor.device
    rng = _check_random_state(random_state)
    if isinstance(init, str) and init == "random":
        glob = np.random.mtrand._rand
        factors = [torch.tensor(glob.random_sample((s, rank)), dtype=dt, device=dev) for s in tensor.shape]
    elif isinstance(init, str) and init == "svd":
        from ..base import unfold
        from ..svd import svd_interface
        factors = []
        for mode in range(n_modes):
            U, S, _ = svd_interface(unfold(tensor, mode), n_eigenvecs=rank, method=svd)
            if mode == 0:
                k = min(rank, S.shape[0])
                U = U.clone()
                U[:, :k] = U[:, :k] * S[:k]
            if tensor.shape[mode] < rank:
                pad = torch.tensor(rng.random_sample((U.shape[0], rank - tensor.shape[mode])), dtype=dt, device=dev)
                U = torch.cat([U, pad], dim=1)
            factors.append(U[:, :rank].contiguous())
    elif isinstance(init, (tuple, list, CPTensor)):
        try:
            w, f = CPTensor(init)
            f = [x.to(dt).clone() for x in f]
            if bool(torch.all(w == 1)):
                return CPTensor((None, f))
            avg = torch.prod(w) ** (1.0 / w.shape[0])
            return CPTensor((None, [x * avg for x in f]))
        except ValueError:
            raise ValueError("If initialization method is a mapping, then it must "
                             "be possible to convert it to a CPTensor instance")
    else:
        raise ValueError(f'Initialization method "{init}" not recognized')
    factors = [proximal_operator(f, n_const=n_modes, order=i, **cons) for i, f in enumerate(factors)]
    return CPTensor((None, factors))


def constrained_parafac(tensor, rank, n_iter_max=100, n_iter_max_inner=10, init="svd", svd="truncated_svd",
                        tol_outer=1e-8, tol_inner=1e-6, random_state=None, verbose=0, return_errors=False,
                        cvg_criterion="abs_rec_error", fixed_modes=None, non_negative=None, l1_reg=None, l2_reg=None,
                        l2_square_reg=None, unimodality=None, normalize=None, simplex=None, normalized_sparsity=None,
                        soft_sparsity=None, smoothness=None, monotonicity=None, hard_sparsity=None):
    """CP decomposition with per-mode constraints or penalties by AO-ADMM
    (``_constrained_cp.py:166-424``); ``tensor`` is a float32 / float64 GPU
    tensor, rank <= 64.  Returns ``CPTensor`` (and the per-iteration
    relative errors with ``return_errors``)."""
    cons = dict(non_negative=non_negative, l1_reg=l1_reg, l2_reg=l2_reg, l2_square_reg=l2_square_reg,
                unimodality=unimodality, normalize=normalize, simplex=simplex,
                normalized_sparsity=normalized_sparsity, soft_sparsity=soft_sparsity, smoothness=smoothness,
                monotonicity=monotonicity, hard_sparsity=hard_sparsity)
    _lib.require_gpu_tensor(tensor, "tensor")
    if tensor.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"constrained_parafac takes float32 / float64 tensors, got {tensor.dtype}")
    tensor = tensor.contiguous()
    nd = tensor.dim()
    rank = validate_cp_rank(tuple(tensor.shape), rank=rank)
    validate_constraints(n_const=nd, **cons)
    weights, factors = initialize_constrained_parafac(tensor, rank, init=init, svd=svd, random_state=random_state,
                                                      **cons)
    dev, dt = tensor.device, tensor.dtype
    if weights is None:
        weights = torch.ones(rank, dtype=dt, device=dev)
    factors = [f.to(dt).contiguous().clone() for f in factors]
    normx2 = _sqnorm64(tensor).reshape(1)

    fixed_modes = [] if fixed_modes is None else list(fixed_modes)
    if nd - 1 in fixed_modes:
        warnings.warn("You asked for fixing the last mode, which is not supported.\n "
                      "The last mode will not be fixed. Consider using tl.moveaxis()")
        fixed_modes.remove(nd - 1)
    modes_list = [m for m in range(nd) if m not in fixed_modes]
    codes = [_code_and_param(*validate_constraints(n_const=nd, order=m, **cons)) for m in range(nd)]

    duals = [torch.zeros(f.shape, dtype=torch.float64, device=dev) for f in factors]
    aux = [torch.zeros(f.shape, dtype=torch.float64, device=dev) for f in factors]  # x_split^T
    grams = grams64(factors)
    status = torch.zeros(nd, dtype=torch.int32, device=dev)
    norms = torch.zeros((nd, 2), dtype=torch.float64, device=dev)
    iprod = torch.zeros(1, dtype=torch.float64, device=dev)
    errs_dev = torch.zeros(max(1, n_iter_max), dtype=torch.float64, device=dev)
    gp, _keep_g = ptr_array(grams)
    rec_errors = []
    n_done = 0
    for iteration in range(n_iter_max):
        if verbose > 1:
            print("Starting iteration", iteration + 1)
        for mode in modes_list:
            if verbose > 1:
                print("Mode", mode, "of", nd)
            mttkrp = unfolding_dot_khatri_rao(tensor, (None, factors), mode)
            code, p = codes[mode]
            admm_update(factors[mode], aux[mode], duals[mode], mttkrp, code, p, n_iter_max_inner, tol_inner,
                        grams=grams, mode=mode, status=status[mode:mode + 1], norms=norms[mode],
                        iprod=iprod if mode == nd - 1 else None)
            _gram_into(factors[mode], grams[mode])
        # error from the Grams and the last mode's sum(UtM * x) (weights are ones)
        check(lib().tla_cp_error(gp, nd, None, _lib.dtype_code(factors[0]), rank, ptr(normx2), ptr(iprod),
                                 ptr(errs_dev), iteration, stream(dev)), "cp_error")
        n_done = iteration + 1
        if tol_outer:
            if int(torch.max(status)) == 2:
                raise np.linalg.LinAlgError("Singular matrix")
            rec_errors = [float(e) for e in errs_dev[:n_done].cpu()]
            cn = norms[modes_list].cpu().numpy()
            constraint_error = float(np.sum(cn[:, 0] / cn[:, 1]))
            if iteration >= 1:
                dec = rec_errors[-2] - rec_errors[-1]
                if verbose:
                    print(f"iteration {iteration}, reconstruction error: {rec_errors[-1]}, decrease = {dec}")
                if constraint_error < tol_outer:
                    break
                if cvg_criterion == "abs_rec_error":
                    stop = abs(dec) < tol_outer
                elif cvg_criterion == "rec_error":
                    stop = dec < tol_outer
                else:
                    raise TypeError("Unknown convergence criterion")
                if stop:
                    if verbose:
                        print(f"PARAFAC converged after {iteration} iterations")
                    break
            elif verbose:
                print(f"reconstruction error={rec_errors[-1]}")
    if int(torch.max(status)) == 2:
        raise np.linalg.LinAlgError("Singular matrix")
    cp_tensor = CPTensor((weights, factors))
    if return_errors:
        return cp_tensor, [float(e) for e in errs_dev[:n_done].cpu()]
    return cp_tensor


def _gram_into(factor, G):
    rows, rank = factor.shape
    nb = lib().tla_gram64_workspace_size(rows, rank)
    ws = _lib.workspace(nb, factor.device)
    check(lib().tla_gram64(_lib.dtype_code(factor), ptr(factor), rows, rank, ptr(G), ptr(ws), nb,
                           stream(factor.device)), "gram64")


class ConstrainedCP:
    """Scikit-style wrapper (``_constrained_cp.py:427-``):
    ``ConstrainedCP(rank, ...).fit_transform(X)``."""

    def __init__(self, rank, n_iter_max=100, n_iter_max_inner=10, init="svd", svd="truncated_svd", tol_outer=1e-8,
                 tol_inner=1e-6, random_state=None, verbose=0, cvg_criterion="abs_rec_error", fixed_modes=None,
                 **constraints):
        self.rank, self.n_iter_max, self.n_iter_max_inner = rank, n_iter_max, n_iter_max_inner
        self.init, self.svd, self.tol_outer, self.tol_inner = init, svd, tol_outer, tol_inner
        self.random_state, self.verbose, self.cvg_criterion = random_state, verbose, cvg_criterion
        self.fixed_modes = fixed_modes
        unknown = set(constraints) - set(_CON_KW)
        if unknown:
            raise TypeError(f"unknown constraints {sorted(unknown)}")
        self.constraints = constraints

    def fit_transform(self, tensor):
        cp, errors = constrained_parafac(tensor, self.rank, n_iter_max=self.n_iter_max,
                                         n_iter_max_inner=self.n_iter_max_inner, init=self.init, svd=self.svd,
                                         tol_outer=self.tol_outer, tol_inner=self.tol_inner,
                                         random_state=self.random_state, verbose=self.verbose, return_errors=True,
                                         cvg_criterion=self.cvg_criterion, fixed_modes=self.fixed_modes,
                                         **self.constraints)
        self.decomposition_ = cp
        self.errors_ = errors
        return cp

    def fit(self, tensor):
        self.fit_transform(tensor)
        return self

    def __repr__(self):
        return f"Rank-{self.rank} Constrained CP decomposition."
