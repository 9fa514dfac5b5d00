"""CP tensor type and operations on GPU tensors.

Follows tensorly/cp_tensor.py (0.9.0): :class:`CPTensor` (``:17-160``),
:func:`validate_cp_tensor` (``:163-214``), :func:`validate_cp_rank`
(``:240-275``), :func:`cp_normalize` (``:278-321``), :func:`cp_to_tensor`
(``:433-485``), :func:`cp_norm` (``:614-644``).
"""
from collections.abc import Mapping

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, ptr_array, stream, workspace
from .tenalg import khatri_rao

__all__ = ["CPTensor", "validate_cp_tensor", "validate_cp_rank", "cp_normalize", "cp_to_tensor",
           "cp_norm", "grams64", "cp_lstsq_grad"]


class CPTensor(Mapping):
    """``(weights, factors)`` pair, indexable as ``cp[0]``/``cp[1]`` and
    unpackable like the reference type."""

    def __init__(self, cp_tensor):
        super().__init__()
        shape, rank = validate_cp_tensor(cp_tensor)
        weights, factors = cp_tensor
        if weights is None:
            weights = torch.ones(rank, dtype=factors[0].dtype, device=factors[0].device)
        self.shape = shape
        self.rank = rank
        self.factors = factors
        self.weights = weights

    def __getitem__(self, index):
        if index == 0:
            return self.weights
        if index == 1:
            return self.factors
        raise IndexError(f"You tried to access index {index} of a CP tensor.\n"
                         "You can only access index 0 and 1 of a CP tensor"
                         "(corresponding respectively to the weights and factors)")

    def __setitem__(self, index, value):
        if index == 0:
            self.weights = value
        elif index == 1:
            self.factors = value
        else:
            raise IndexError(f"You tried to set the value at index {index} of a CP tensor.")

    def __iter__(self):
        yield self.weights
        yield self.factors

    def __len__(self):
        return 2

    def __repr__(self):
        return f"(weights, factors) : rank-{self.rank} CPTensor of shape {self.shape}"

    def to_tensor(self):
        return cp_to_tensor(self)

    def norm(self):
        return cp_norm(self)

    def normalize(self, inplace=True):
        w, f = cp_normalize(self)
        if inplace:
            self.weights, self.factors = w, f
            return self
        return CPTensor((w, f))

    def cp_copy(self):
        return CPTensor((self.weights.clone(), [f.clone() for f in self.factors]))


def validate_cp_tensor(cp_tensor):
    """Returns ``(shape, rank)``; ValueError on inconsistent factors
    (``cp_tensor.py:163-214``)."""
    if isinstance(cp_tensor, CPTensor):
        return cp_tensor.shape, cp_tensor.rank
    weights, factors = cp_tensor
    if factors[0].dim() == 2:
        rank = int(factors[0].shape[1])
    elif factors[0].dim() == 1:
        rank = 1
    else:
        raise ValueError("Got a factor with 3 dimensions but CP factors should be at most 2D, "
                         "of shape (size, rank).")
    shape = []
    for i, f in enumerate(factors):
        size, r = (tuple(f.shape) if f.dim() == 2 else (f.shape[0], 1))
        if r != rank:
            raise ValueError("All the factors of a CP tensor should have the same number of column."
                             f"However, factors[0].shape[1]={rank} but factors[{i}].shape[1]={r}.")
        shape.append(int(size))
    if weights is not None and tuple(weights.shape) != (rank,):
        raise ValueError(f"Given factors for a rank-{rank} CP tensor but len(weights)={tuple(weights.shape)}.")
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


def grams64(factors):
    """FP64 Grams ``A^T A`` of each factor on the GPU (one kernel each)."""
    out = []
    for f in factors:
        f = f.contiguous()
        rows, rank = f.shape
        g = torch.empty((rank, rank), dtype=torch.float64, device=f.device)
        nb = lib().tla_gram64_workspace_size(rows, rank)
        ws = workspace(nb, f.device)
        check(lib().tla_gram64(_lib.dtype_code(f), ptr(f), rows, rank, ptr(g), ptr(ws), nb, stream(f.device)),
              "gram64")
        out.append(g)
    return out


def cp_norm(cp_tensor):
    """``sqrt(sum((*_m A_m^T A_m) * w w^T))`` (``cp_tensor.py:614-644``)
    from FP64 Grams; returns a 0-d tensor of the factors' dtype."""
    _, rank = validate_cp_tensor(cp_tensor)
    weights, factors = cp_tensor
    for i, f in enumerate(factors):
        _lib.require_gpu_tensor(f, f"factors[{i}]")
    gs = grams64([f.reshape(f.shape[0], rank) for f in factors])
    acc = gs[0]
    for g in gs[1:]:
        acc = acc * g
    if weights is not None:
        w = weights.to(torch.float64)
        acc = acc * (w[:, None] * w[None, :])
    return torch.sqrt(acc.sum()).to(factors[0].dtype)


def cp_normalize(cp_tensor):
    """Weights folded into the first factor, every column scaled to unit
    l2 norm, norms collected in the weights (``cp_tensor.py:278-321``)."""
    _, rank = validate_cp_tensor(cp_tensor)
    weights, factors = cp_tensor
    if weights is None:
        weights = torch.ones(rank, dtype=factors[0].dtype, device=factors[0].device)
    out = []
    for i, f in enumerate(factors):
        if i == 0:
            f = f * weights
            weights = torch.ones(rank, dtype=f.dtype, device=f.device)
        scales = torch.linalg.vector_norm(f, dim=0)
        safe = torch.where(scales == 0, torch.ones_like(scales), scales)
        weights = weights * scales
        out.append(f / safe.reshape(1, -1))
    return CPTensor((weights, out))


def cp_to_tensor(cp_tensor, mask=None):
    """Dense tensor ``fold((A_0 * w) @ khatri_rao(A_1..)^T, 0)``
    (``cp_tensor.py:433-485``); the Khatri-Rao product comes from the HIP
    kernel, the product is one library GEMM."""
    shape, _ = validate_cp_tensor(cp_tensor)
    weights, factors = cp_tensor
    if len(shape) == 1:
        return torch.sum(weights * factors[0], dim=1)
    if weights is None:
        weights = 1
    if mask is None:
        full = (factors[0] * weights) @ khatri_rao(factors, skip_matrix=0).T
    else:
        full = torch.sum(khatri_rao([factors[0] * weights] + list(factors[1:]), mask=mask), dim=1)
    return full.reshape(shape)


def cp_lstsq_grad(cp_tensor, tensor, return_loss=False, mask=None):
    """Gradient of ``0.5 ||X - [[w; A]]||^2`` with respect to each factor
    (``cp_tensor.py:377-430``; not with respect to the weights), returned as
    ``CPTensor((None, grads))``, and the loss when ``return_loss``.

    The reference forms the residual tensor and runs one MTTKRP of it per
    mode.  Here ``-MTTKRP(X - M, n) = A_n accum_n - MTTKRP(X, n)`` with
    ``accum_n`` = (Hadamard of the other modes' Grams) * w w^T, so each mode
    is one fused MTTKRP pass over X and one ``tla_cp_lstsq_grad_mode``
    launch, and the loss is ``0.5 (||X||^2 + ||M||^2 - 2 <X, M>)`` in FP64
    from the Grams and the last mode's MTTKRP.  No reconstruction is formed.
    Near an exact fit both forms cancel; this one loses about
    ``||X|| / ||X - M||`` relative digits more.  ``mask`` raises
    NotImplementedError."""
    from .tenalg import unfolding_dot_khatri_rao
    if mask is not None:
        raise NotImplementedError("mask is not supported by tensorly_amd.cp_lstsq_grad")
    _lib.require_gpu_tensor(tensor, "tensor")
    _, rank = validate_cp_tensor(cp_tensor)
    weights, factors = cp_tensor
    tensor = tensor.contiguous()
    dt = tensor.dtype
    if dt not in (torch.float32, torch.float64):
        raise TypeError(f"cp_lstsq_grad takes float32 / float64 tensors, got {dt}")
    factors = [f.to(dt).contiguous() for f in factors]
    w = (torch.ones(rank, dtype=dt, device=tensor.device) if weights is None
         else weights.to(dt).contiguous())
    grams = grams64(factors)
    gp, _keep = ptr_array(grams)
    grads = []
    m = None
    for n, f in enumerate(factors):
        m = unfolding_dot_khatri_rao(tensor, (w, factors), n).contiguous()
        g = torch.empty_like(f)
        check(lib().tla_cp_lstsq_grad_mode(_lib.dtype_code(f), ptr(f), ptr(m), ptr(g), f.shape[0], rank, gp,
                                           len(factors), n, ptr(w), _lib.dtype_code(w), stream(f.device)),
              "cp_lstsq_grad")
        grads.append(g)
    grad = CPTensor((None, grads))
    if not return_loss:
        return grad
    nx2 = _sqnorm(tensor)
    acc = grams[0]
    for gm in grams[1:]:
        acc = acc * gm
    w64 = w.to(torch.float64)
    nm2 = torch.sum(acc * (w64[:, None] * w64[None, :]))
    ip = torch.sum(m.to(torch.float64) * factors[-1].to(torch.float64))
    return grad, 0.5 * (nx2 + nm2 - 2.0 * ip)


def _sqnorm(x):
    out = torch.empty(1, dtype=torch.float64, device=x.device)
    nb = lib().tla_sqnorm_workspace_size(x.numel())
    ws = workspace(nb, x.device)
    check(lib().tla_sqnorm(_lib.dtype_code(x), ptr(x), x.numel(), ptr(out), ptr(ws), nb, stream(x.device)), "sqnorm")
    return out[0]
