"""Proximal operators and the ADMM solver of constrained CP on the gfx950
kernels (``tensorly/tenalg/proximal.py`` and ``tensorly/solvers/admm.py``,
0.9.0).

:func:`validate_constraints` is host logic (which constraint and parameter a
mode gets, ``proximal.py:4-150``); :func:`proximal_operator` and the named
operators run ``tla_prox`` (one workgroup, FP64 arithmetic) on a GPU matrix
of at most 64 columns; :func:`admm` runs the whole inner loop of
``admm.py:5-155`` in one ``tla_admm_update`` launch.
"""
import torch

from . import _lib
from ._lib import check, lib, ptr, stream

__all__ = ["CONSTRAINTS", "validate_constraints", "proximal_operator", "soft_thresholding", "l2_prox",
           "l2_square_prox", "hard_thresholding", "normalized_sparsity_prox", "simplex_prox", "soft_sparsity_prox",
           "smoothness_prox", "monotonicity_prox", "unimodality_prox", "admm"]

# constraint keyword -> C-ABI code, in the reference's order (the order in
# which later keywords overwrite earlier ones, proximal.py:84-146)
CONSTRAINTS = {
    "non_negative": 1, "l1_reg": 2, "l2_reg": 3, "l2_square_reg": 4, "unimodality": 5, "normalize": 6,
    "simplex": 7, "normalized_sparsity": 8, "soft_sparsity": 9, "smoothness": 10, "monotonicity": 11,
    "hard_sparsity": 12,
}


def validate_constraints(non_negative=None, l1_reg=None, l2_reg=None, l2_square_reg=None, unimodality=None,
                         normalize=None, simplex=None, normalized_sparsity=None, soft_sparsity=None,
                         smoothness=None, monotonicity=None, hard_sparsity=None, n_const=1, order=0):
    """Constraint name and parameter of mode ``order`` (``proximal.py:4-150``).

    Each keyword is unused (None/False), a scalar or True (every mode), a
    list (one entry per mode) or a dict (mode -> parameter).  A mode
    constrained twice raises ValueError.  As in the reference, a list
    registers its name for every mode (its entry as the parameter) and later
    keywords overwrite earlier ones."""
    given = dict(non_negative=non_negative, l1_reg=l1_reg, l2_reg=l2_reg, l2_square_reg=l2_square_reg,
                 unimodality=unimodality, normalize=normalize, simplex=simplex,
                 normalized_sparsity=normalized_sparsity, soft_sparsity=soft_sparsity, smoothness=smoothness,
                 monotonicity=monotonicity, hard_sparsity=hard_sparsity)
    taken = set()
    for name in CONSTRAINTS:
        spec = given[name]
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
    for name in CONSTRAINTS:
        spec = given[name]
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


def _code_and_param(name, param):
    if name is None:
        return 0, 0.0
    code = CONSTRAINTS[name]
    needs_value = name in ("l1_reg", "l2_reg", "l2_square_reg", "simplex", "normalized_sparsity", "soft_sparsity",
                           "smoothness", "hard_sparsity")
    if needs_value and (param is None or isinstance(param, bool)):
        raise TypeError(f"{name} needs a numeric parameter, got {param!r}")
    return code, float(param) if needs_value else 0.0


def _prox(tensor, code, param):
    _lib.require_gpu_tensor(tensor, "tensor")
    if tensor.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"proximal operators take float32 / float64 tensors, got {tensor.dtype}")
    one_d = tensor.dim() == 1
    if tensor.dim() > 2:
        raise ValueError("The proximal operators take a matrix (or a vector).")
    x = tensor.reshape(tensor.shape[0], -1).contiguous()
    rows, cols = x.shape
    out = torch.empty_like(x)
    nb = lib().tla_prox_workspace_size(rows, cols)
    ws = _lib.workspace(nb, x.device)
    check(lib().tla_prox(_lib.dtype_code(x), ptr(x), ptr(out), rows, cols, code, float(param), ptr(ws), nb,
                         stream(x.device)), "prox")
    return out.reshape(-1) if one_d else out


def proximal_operator(tensor, non_negative=None, l1_reg=None, l2_reg=None, l2_square_reg=None, unimodality=None,
                      normalize=None, simplex=None, normalized_sparsity=None, soft_sparsity=None, smoothness=None,
                      monotonicity=None, hard_sparsity=None, n_const=1, order=0):
    """``proximal.py:153-267``: the proximal operator of mode ``order``'s
    constraint applied to ``tensor`` (unchanged when none, or when
    ``n_const`` is None)."""
    if n_const is None:
        return tensor
    name, param = validate_constraints(non_negative=non_negative, l1_reg=l1_reg, l2_reg=l2_reg,
                                       l2_square_reg=l2_square_reg, unimodality=unimodality, normalize=normalize,
                                       simplex=simplex, normalized_sparsity=normalized_sparsity,
                                       soft_sparsity=soft_sparsity, smoothness=smoothness,
                                       monotonicity=monotonicity, hard_sparsity=hard_sparsity, n_const=n_const,
                                       order=order)
    if name is None:
        return tensor
    code, p = _code_and_param(name, param)
    return _prox(tensor, code, p)


def soft_thresholding(tensor, threshold):
    """``proximal.py:659-689`` (scalar threshold)."""
    return _prox(tensor, CONSTRAINTS["l1_reg"], float(threshold))


def l2_prox(tensor, regularizer):
    """``proximal.py:503-529``."""
    return _prox(tensor, CONSTRAINTS["l2_reg"], float(regularizer))


def l2_square_prox(tensor, regularizer):
    """``proximal.py:478-500``."""
    return _prox(tensor, CONSTRAINTS["l2_square_reg"], float(regularizer))


def hard_thresholding(tensor, number_of_non_zero):
    """``proximal.py:634-656``."""
    return _prox(tensor, CONSTRAINTS["hard_sparsity"], float(number_of_non_zero))


def normalized_sparsity_prox(tensor, threshold):
    """``proximal.py:532-561``."""
    return _prox(tensor, CONSTRAINTS["normalized_sparsity"], float(threshold))


def simplex_prox(tensor, parameter):
    """``proximal.py:592-631``."""
    return _prox(tensor, CONSTRAINTS["simplex"], float(parameter))


def soft_sparsity_prox(tensor, threshold):
    """``proximal.py:564-589``."""
    return _prox(tensor, CONSTRAINTS["soft_sparsity"], float(threshold))


def smoothness_prox(tensor, regularizer):
    """``proximal.py:270-291``."""
    return _prox(tensor, CONSTRAINTS["smoothness"], float(regularizer))


def monotonicity_prox(tensor, decreasing=False):
    """``proximal.py:294-356`` (non-decreasing; ``decreasing`` flips the
    rows around the operator)."""
    if decreasing:
        return torch.flip(_prox(torch.flip(tensor, (0,)), CONSTRAINTS["monotonicity"], 0.0), (0,))
    return _prox(tensor, CONSTRAINTS["monotonicity"], 0.0)


def unimodality_prox(tensor):
    """``proximal.py:359-475``."""
    return _prox(tensor, CONSTRAINTS["unimodality"], 0.0)


def admm_update(factor, x_split, dual, mttkrp, code, param, n_iter_max, tol, utu=None, grams=None, mode=0,
                status=None, iters=None, norms=None, iprod=None):
    """One ``tla_admm_update`` launch: ``factor`` (rows x R, float32 /
    float64) updated in place, ``x_split`` (x_split^T, rows x R FP64)
    written, ``dual`` (rows x R FP64) updated; optional device outputs."""
    rows, rank = factor.shape
    dev = factor.device
    nb = lib().tla_admm_workspace_size(rows, rank)
    ws = _lib.workspace(nb, dev)
    gp, _keep = _lib.ptr_array(grams) if grams is not None else (None, None)
    check(lib().tla_admm_update(_lib.dtype_code(factor), ptr(factor), ptr(x_split), ptr(dual),
                                ptr(mttkrp.contiguous()), rows, rank, ptr(utu), gp, len(grams) if grams else 0, mode,
                                int(n_iter_max), float(tol), int(code), float(param), ptr(status), ptr(iters),
                                ptr(norms), ptr(iprod), ptr(ws), nb, stream(dev)),
          "admm_update")


def admm(UtM, UtU, x, dual_var, n_iter_max=100, n_const=None, order=None, non_negative=None, l1_reg=None,
         l2_reg=None, l2_square_reg=None, unimodality=None, normalize=None, simplex=None, normalized_sparsity=None,
         soft_sparsity=None, smoothness=None, monotonicity=None, hard_sparsity=None, tol=1e-4):
    """``solvers/admm.py:5-155``: returns ``(x, x_split, dual_var)`` (new
    tensors; ``x_split`` is R x rows as in the reference).  ``UtM`` is rows x
    R (the MTTKRP), ``UtU`` R x R, ``x`` and ``dual_var`` rows x R.  Raises
    ``numpy.linalg.LinAlgError`` when UtU + rho I is singular."""
    import numpy as np
    if n_const is None:
        raise NotImplementedError("admm without constraints (n_const=None) is a plain least-squares solve; "
                                  "use parafac")
    _lib.require_gpu_tensor(x, "x")
    dev = x.device
    for t, nm in ((UtM, "UtM"), (UtU, "UtU"), (dual_var, "dual_var")):
        _lib.require_on_device(t, dev, nm)
    name, param = validate_constraints(non_negative=non_negative, l1_reg=l1_reg, l2_reg=l2_reg,
                                       l2_square_reg=l2_square_reg, unimodality=unimodality, normalize=normalize,
                                       simplex=simplex, normalized_sparsity=normalized_sparsity,
                                       soft_sparsity=soft_sparsity, smoothness=smoothness,
                                       monotonicity=monotonicity, hard_sparsity=hard_sparsity, n_const=n_const,
                                       order=order)
    code, p = _code_and_param(name, param)
    xo = x.clone().contiguous()
    d = dual_var.to(torch.float64).clone().contiguous()
    xs = torch.empty(xo.shape, dtype=torch.float64, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    admm_update(xo, xs, d, UtM.to(xo.dtype), code, p, n_iter_max, tol, utu=UtU.to(torch.float64).contiguous(),
                status=status)
    if int(status) == 2:
        raise np.linalg.LinAlgError("Singular matrix")
    return xo, xs.t(), d
