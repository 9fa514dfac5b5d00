"""Tucker / HOOI drop-in (``tensorly/decomposition/_tucker.py``, 0.9.0).

:func:`tucker` (``:224-345``), :func:`partial_tucker` (``:105-221``) and
:func:`initialize_tucker` (``:25-102``) keep the reference signatures.  The
TTM chains run on the MFMA mode_dot kernels; the factor update
(``:194-202``: leading left singular vectors of the partially projected
tensor's unfolding) is the FP64 Gram eigensolve, warm-started from the
previous factor, with the ``svd_flip`` sign rule; ``‖X‖`` and ``‖core‖`` are
FP64 reductions.  ``mask`` is not supported (NotImplementedError).

Extension (keyword-only): ``precision`` for the TTM arithmetic.
"""
import math
import warnings

import numpy as np
import torch

from .. import _lib
from ..svd import leading_left_singular_vectors, leading_left_singular_vectors_batch
from ..tenalg import mode_dot, multi_mode_dot, precision as precision_ctx, get_precision
from ..tucker_tensor import TuckerTensor, validate_tucker_rank

__all__ = ["tucker", "partial_tucker", "initialize_tucker", "Tucker"]


def _sqnorm64(x):
    x = x.contiguous()
    out = torch.empty(1, dtype=torch.float64, device=x.device)
    L = _lib.lib()
    nb = L.tla_sqnorm_workspace_size(x.numel())
    ws = _lib.workspace(nb, x.device)
    _lib.check(L.tla_sqnorm(_lib.dtype_code(x), _lib.ptr(x), x.numel(), _lib.ptr(out), _lib.ptr(ws), nb,
                            _lib.stream(x.device)), "sqnorm")
    return out[0]


def initialize_tucker(tensor, rank, modes, random_state, init="svd", svd="truncated_svd", non_negative=False,
                      mask=None, svd_mask_repeats=5, warm64=None):
    """``_tucker.py:25-102`` (no mask).  With init='svd' the modes' SVDs are
    independent: their Gram eigensolves run side by side in one launch.
    ``warm64`` (a list, optional) receives the FP64 factors of the SVD
    init."""
    if mask is not None:
        raise NotImplementedError("mask is not supported by tensorly_amd.tucker")
    if isinstance(init, str) and init == "svd":
        if svd not in ("truncated_svd", "symeig_svd"):
            raise ValueError(f"Got svd={svd}. However, the possible choices are ['truncated_svd', 'symeig_svd'].")
        res = leading_left_singular_vectors_batch(tensor, list(modes), [rank[i] for i in range(len(modes))])
        factors = [r[0] for r in res]
        if warm64 is not None:
            warm64.extend(r[2] for r in res)
        core = multi_mode_dot(tensor, factors, modes=modes, transpose=True)
    elif isinstance(init, str) and init == "random":
        from ._cp import _check_random_state
        rng = _check_random_state(random_state)
        dt, dev = tensor.dtype, tensor.device
        core = torch.tensor(rng.random_sample([rank[i] for i in range(len(modes))]) + 0.01, dtype=dt, device=dev)
        factors = [torch.tensor(rng.random_sample((tensor.shape[m], rank[i])), dtype=dt, device=dev)
                   for i, m in enumerate(modes)]
    else:
        core, factors = init
        factors = list(factors)
    if non_negative is True:
        factors = [torch.abs(f) for f in factors]
        core = torch.abs(core)
    return core, factors


def partial_tucker(tensor, rank, modes=None, n_iter_max=100, init="svd", tol=10e-5, svd="truncated_svd",
                   random_state=None, verbose=False, mask=None, svd_mask_repeats=5, *, precision=None):
    """HOOI on ``modes`` (``_tucker.py:105-221``).  Returns
    ``((core, factors), rec_errors)``."""
    _lib.require_gpu_tensor(tensor, "tensor")
    if mask is not None:
        raise NotImplementedError("mask is not supported by tensorly_amd.tucker")
    if modes is None:
        modes = list(range(tensor.dim()))
    if rank is None:
        warnings.warn("No value given for 'rank'. The decomposition will preserve the original size.", Warning)
        rank = [tensor.shape[m] for m in modes]
    elif isinstance(rank, int):
        warnings.warn(f"Given only one int for 'rank' instead of a list of {len(modes)} modes. "
                      "Using this rank for all modes.", Warning)
        rank = tuple(rank for _ in modes)
    else:
        rank = tuple(rank)
    with precision_ctx(precision or get_precision()):
        # FP64 warm starts for the next eigensolve of each mode: the init's
        # FP64 vectors, then the solver's whole final block (accurate past
        # float32, and its guard columns need no re-orthonormalising)
        warm64 = []
        core, factors = initialize_tucker(tensor, rank, modes, init=init, svd=svd, random_state=random_state,
                                          warm64=warm64)
        if len(warm64) != len(modes):
            warm64 = [None] * len(modes)
        nx2 = float(_sqnorm64(tensor))
        norm_tensor = math.sqrt(nx2)

        def rec_error(c2):  # the reference's expression, in FP64 on the host
            return math.sqrt(abs(nx2 - c2)) / norm_tensor

        # with the modes in ascending order the last mode's projected tensor
        # is the core's chain without its last product (multi_mode_dot
        # applies the modes sorted), so the core is one small product away
        sorted_modes = list(modes) == sorted(set(modes))
        # and the products of the tensor with the first index modes' new
        # factors (prefix) are shared by every later mode of the iteration:
        # y_index = prefix_index x (the later modes), the same sequence of
        # products as multi_mode_dot(..., skip=index)
        core_sq = []  # ||core||^2 per iteration, FP64 device scalars
        for iteration in range(n_iter_max):
            prefix = tensor
            for index, mode in enumerate(modes):
                if sorted_modes:
                    y = prefix
                    for later in range(index + 1, len(modes)):
                        y = mode_dot(y, factors[later], modes[later], transpose=True)
                else:
                    y = multi_mode_dot(tensor, factors, modes=modes, skip=index, transpose=True)
                U, _, warm64[index] = leading_left_singular_vectors(y, mode, rank[index], warm=warm64[index],
                                                                    return_warm=True)
                factors[index] = U
                if sorted_modes and index + 1 < len(modes):
                    prefix = mode_dot(prefix, U, mode, transpose=True)
            if sorted_modes and modes:
                core = mode_dot(y, factors[-1], modes[-1], transpose=True)
            else:
                core = multi_mode_dot(tensor, factors, modes=modes, transpose=True)
            core_sq.append(_sqnorm64(core))
            if iteration > 1 and (tol or verbose):
                e1, e2 = rec_error(float(core_sq[-2])), rec_error(float(core_sq[-1]))
                if verbose:
                    print(f"reconstruction error={e2}, variation={e1 - e2}.")
                if tol and abs(e1 - e2) < tol:
                    if verbose:
                        print(f"converged in {iteration} iterations.")
                    break
        rec_errors = [rec_error(float(c2)) for c2 in torch.stack(core_sq).cpu()] if core_sq else []
    return (core, factors), rec_errors


def tucker(tensor, rank, fixed_factors=None, n_iter_max=100, init="svd", return_errors=False,
           svd="truncated_svd", tol=10e-5, random_state=None, mask=None, verbose=False, *, precision=None):
    """Tucker decomposition by HOOI (``_tucker.py:224-345``)."""
    if fixed_factors:
        try:
            core, factors = init
        except Exception:
            raise ValueError(f'Got fixed_factor={fixed_factors} but no appropriate Tucker tensor was passed for "init".')
        fixed_factors = sorted(fixed_factors)
        fixed = [(i, f) for i, f in enumerate(factors) if i in fixed_factors]
        modes_fixed = [i for i, _ in fixed]
        factors_fixed = [f for _, f in fixed]
        core = multi_mode_dot(core, factors_fixed, modes=modes_fixed)
        free = [(i, f) for i, f in enumerate(factors) if i not in fixed_factors]
        modes = [i for i, _ in free]
        (core, new_factors), rec_errors = partial_tucker(
            tensor, rank=rank, modes=modes, n_iter_max=n_iter_max, init=(core, [f for _, f in free]), svd=svd,
            tol=tol, random_state=random_state, mask=mask, verbose=verbose, precision=precision)
        factors = list(new_factors)
        for i, e in enumerate(fixed_factors):
            factors.insert(e, factors_fixed[i])
        core = multi_mode_dot(core, factors_fixed, modes=modes_fixed, transpose=True)
        return TuckerTensor((core, factors))
    modes = list(range(tensor.dim()))
    rank = validate_tucker_rank(tuple(tensor.shape), rank=rank)
    (core, factors), rec_errors = partial_tucker(tensor, rank=rank, modes=modes, n_iter_max=n_iter_max, init=init,
                                                 svd=svd, tol=tol, random_state=random_state, mask=mask,
                                                 verbose=verbose, precision=precision)
    tt = TuckerTensor((core, factors))
    if return_errors:
        return tt, rec_errors
    return tt


class Tucker:
    """Scikit-style wrapper (``_tucker.py:697-797``)."""

    def __init__(self, rank=None, n_iter_max=100, init="svd", return_errors=False, svd="truncated_svd", tol=10e-5,
                 fixed_factors=None, random_state=None, mask=None, verbose=False):
        self.rank, self.n_iter_max, self.init, self.return_errors = rank, n_iter_max, init, return_errors
        self.svd, self.tol, self.fixed_factors, self.random_state = svd, tol, fixed_factors, random_state
        self.mask, self.verbose = mask, verbose

    def fit_transform(self, tensor):
        res = tucker(tensor, rank=self.rank, fixed_factors=self.fixed_factors, n_iter_max=self.n_iter_max,
                     init=self.init, return_errors=True if not self.fixed_factors else False, svd=self.svd,
                     tol=self.tol, random_state=self.random_state, mask=self.mask, verbose=self.verbose)
        if isinstance(res, tuple):
            self.decomposition_, self.errors_ = res
        else:
            self.decomposition_ = res
        return self.decomposition_

    def fit(self, tensor):
        self.fit_transform(tensor)
        return self

    def __repr__(self):
        return f"Rank-{self.rank} Tucker decomposition via HOOI."
