"""Tucker tensor type (``tensorly/tucker_tensor.py``, 0.9.0): :class:`TuckerTensor`
(``:227-320``), :func:`tucker_to_tensor` (``:50-75``),
:func:`validate_tucker_rank` (``:344-447``)."""
from collections.abc import Mapping
import warnings

import numpy as np

from .tenalg import multi_mode_dot

__all__ = ["TuckerTensor", "tucker_to_tensor", "validate_tucker_rank", "validate_tucker_tensor"]


def validate_tucker_tensor(tucker_tensor):
    core, factors = tucker_tensor
    if len(factors) < 2:
        raise ValueError("A Tucker tensor should be composed of at least two factors and a core."
                         f"However, {len(factors)} factor was given.")
    if len(factors) != core.dim():
        raise ValueError("Tucker decompositions should have one factor per more of the core tensor."
                         f"However, core has {core.dim()} modes but {len(factors)} factors have been provided")
    shape, rank = [], []
    for i, f in enumerate(factors):
        s, r = f.shape
        if r != core.shape[i]:
            raise ValueError("Factor `n` of Tucker decomposition should verify:\n"
                             "factors[n].shape[1] = core.shape[n]."
                             f"However, factors[{i}].shape[1]={r} but core.shape[{i}]={core.shape[i]}.")
        shape.append(int(s))
        rank.append(int(r))
    return tuple(shape), tuple(rank)


class TuckerTensor(Mapping):
    def __init__(self, tucker_tensor):
        super().__init__()
        shape, rank = validate_tucker_tensor(tucker_tensor)
        core, factors = tucker_tensor
        self.shape, self.rank = tuple(shape), tuple(rank)
        self.core, self.factors = core, factors

    def __getitem__(self, index):
        if index == 0:
            return self.core
        if index == 1:
            return self.factors
        raise IndexError(f"You tried to access index {index} of a Tucker tensor.\n"
                         "You can only access index 0 and 1 of a Tucker tensor"
                         "(corresponding respectively to core and factors)")

    def __setitem__(self, index, value):
        if index == 0:
            self.core = value
        elif index == 1:
            self.factors = value
        else:
            raise IndexError(f"You tried to set index {index} of a Tucker tensor.")

    def __iter__(self):
        yield self.core
        yield self.factors

    def __len__(self):
        return 2

    def __repr__(self):
        return f"Decomposed Tucker tensor of shape {self.shape} with rank={self.rank}"

    def to_tensor(self):
        return tucker_to_tensor(self)


def tucker_to_tensor(tucker_tensor, skip_factor=None, transpose_factors=False, modes=None):
    """``multi_mode_dot(core, factors)`` (``tucker_tensor.py:50-75``)."""
    core, factors = tucker_tensor
    return multi_mode_dot(core, factors, skip=skip_factor, transpose=transpose_factors, modes=modes)


def validate_tucker_rank(tensor_shape, rank="same", rounding="round", fixed_modes=None):
    """``tucker_tensor.py:344-447``."""
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
        fixed, n_fixed = [], 0
        if fixed_modes is not None:
            fixed = [(m, shape.pop(m)) for m in sorted(fixed_modes, reverse=True)][::-1]
            n_fixed = np.sum([s ** 2 for _, s in fixed])
            n_modes -= len(fixed)
        sq = np.sum([s ** 2 for s in shape])
        frac = brentq(lambda x: n_param * x ** n_modes + sq * x + n_fixed * x - rank * n_param, 0.0, max(rank, 1.0))
        out = [max(int(fun(s * frac)), 1) for s in shape]
        for m, s in fixed:
            out.insert(m, s)
        return out
    if isinstance(rank, int):
        warnings.warn(f"Given only one int for 'rank' for decomposition a tensor of order {len(tensor_shape)}. "
                      "Using this rank for all modes.", RuntimeWarning)
        if fixed_modes is None:
            return [rank] * len(tensor_shape)
        return [rank if i not in fixed_modes else s for i, s in enumerate(tensor_shape)]
    return rank
