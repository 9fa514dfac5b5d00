"""Matricisation on GPU tensors (``tensorly/base.py:39-79``).

The fused kernels never unfold; these exist for API parity and for the SVD
initialisations.  ``unfold`` of mode 0 is a view, other modes copy (as numpy's
reshape of a moved axis does)."""
import torch

__all__ = ["unfold", "fold"]


def unfold(tensor, mode):
    """``reshape(moveaxis(X, mode, 0), (I_mode, -1))`` (``base.py:39-53``)."""
    return torch.reshape(torch.movedim(tensor, mode, 0), (tensor.shape[mode], -1))


def fold(unfolded_tensor, mode, shape):
    """Inverse of :func:`unfold` (``base.py:56-79``)."""
    full = list(shape)
    lead = full.pop(mode)
    full.insert(0, lead)
    return torch.movedim(torch.reshape(unfolded_tensor, full), 0, mode)
