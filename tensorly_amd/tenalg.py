"""Tenalg drop-ins on the gfx950 kernels.

Signatures, argument semantics and errors follow tensorly's core tenalg
backend (tensorly 0.9.0):

* :func:`khatri_rao`                -- ``tenalg/core_tenalg/_khatri_rao.py:9-109``
* :func:`unfolding_dot_khatri_rao`  -- ``tenalg/core_tenalg/mttkrp.py:9-50``
* :func:`mode_dot`                  -- ``tenalg/core_tenalg/n_mode_product.py:5-76``
* :func:`multi_mode_dot`            -- ``tenalg/core_tenalg/n_mode_product.py:79-135``

Inputs are torch tensors on the GPU (what tensorly's pytorch backend hands
over); outputs are new tensors of the input dtype.  The arithmetic of the
tensor contractions is chosen with :func:`set_precision` /
:func:`precision` ("fp32" default, "bf16", "bf16x3", "valu"); float64
inputs always run in FP64 on the vector ALU.

:func:`register` adds the functions to tensorly's tenalg dispatcher when
tensorly is importable (``tenalg/base_tenalg.py:4-33``,
``tenalg/__init__.py:15-94``).
"""
import contextlib
import threading
import warnings

import torch

from . import _lib
from ._lib import check, lib, ptr, ptr_array, shape_array, stream, workspace

__all__ = ["khatri_rao", "unfolding_dot_khatri_rao", "unfolding_dot_khatri_rao_coo", "mode_dot", "multi_mode_dot",
           "set_precision", "get_precision", "precision", "register", "mttkrp_partial",
           "rankdiag_contract"]

_state = threading.local()


def get_precision():
    return getattr(_state, "precision", "fp32")


def set_precision(name):
    if name not in _lib.PRECISIONS:
        raise ValueError(f"unknown precision {name!r}; choose from {sorted(_lib.PRECISIONS)}")
    _state.precision = name


@contextlib.contextmanager
def precision(name):
    old = get_precision()
    set_precision(name)
    try:
        yield
    finally:
        set_precision(old)


def _prec_code(name=None):
    return _lib.PRECISIONS[name or get_precision()]


def _ndim(t):
    return t.dim()


# ---------------------------------------------------------------------------
def khatri_rao(matrices, weights=None, skip_matrix=None, mask=None):
    """Khatri-Rao product, first matrix varying slowest
    (``_khatri_rao.py:64-109``): ``weights`` scale the first kept matrix, a
    single kept matrix is returned as is, 1-D inputs are single columns."""
    if skip_matrix is not None:
        matrices = [m for i, m in enumerate(matrices) if i != skip_matrix]
    if len(matrices) == 1:
        return matrices[0]
    if _ndim(matrices[0]) == 2:
        ncol = matrices[0].shape[1]
    else:
        ncol = 1
        matrices = [torch.reshape(m, (-1, 1)) for m in matrices]
        warnings.warn("Khatri-rao of a series of vectors instead of matrices. "
                      "Considering each as a matrix with 1 column.")
    for i, m in enumerate(matrices):
        if _ndim(m) != 2:
            raise ValueError("All the matrices must have exactly 2 dimensions!"
                             f"Matrix {i} has dimension {_ndim(m)} != 2.")
        if m.shape[1] != ncol:
            raise ValueError("All matrices must have same number of columns!"
                             f"Matrix {i} has {m.shape[1]} columns != {ncol}.")
    _lib.require_gpu_tensor(matrices[0], "matrices[0]")
    dev = matrices[0].device
    for i, m in enumerate(matrices):
        _lib.require_on_device(m, dev, f"matrices[{i}]")
    if weights is not None:
        _lib.require_on_device(weights, dev, "weights")
    if mask is not None:
        _lib.require_on_device(mask, dev, "mask")
    dt = matrices[0].dtype
    mats = [m.to(dt).contiguous() for m in matrices]
    rows = [m.shape[0] for m in mats]
    n_rows = 1
    for r in rows:
        n_rows *= r
    out = torch.empty((n_rows, ncol), dtype=dt, device=mats[0].device)
    w = weights.to(dt).contiguous() if weights is not None else None
    mk = mask.to(dt).reshape(-1).contiguous() if mask is not None else None
    if len(mats) > 8:
        # fold in chunks of at most 8 factors (kernel limit), keeping order
        head = khatri_rao(mats[:8], weights=weights)
        return khatri_rao([head] + mats[8:], mask=mask)
    pp, _keep = ptr_array(mats)
    sh, _ks = shape_array(rows)
    check(lib().tla_khatri_rao(_lib.dtype_code(mats[0]), pp, len(mats), sh, ncol, ptr(w), ptr(mk),
                               ptr(out), stream(out.device)), "khatri_rao")
    return out


# ---------------------------------------------------------------------------
def _mttkrp_dtype(tensor, factors):
    if tensor.dtype == torch.float64:
        return torch.float64
    return torch.float32


def unfolding_dot_khatri_rao(tensor, cp_tensor, mode):
    """MTTKRP ``unfold(tensor, mode) @ khatri_rao(factors, weights,
    skip_matrix=mode)`` (``mttkrp.py:47-50``), computed by one fused pass over
    the tensor without forming the Khatri-Rao product.

    A bfloat16 ``tensor`` is taken as the BF16 copy of a float32 tensor and
    contracted on BF16 MFMA (float32 factors and output)."""
    weights, factors = cp_tensor
    _lib.require_gpu_tensor(tensor, "tensor")
    nd = _ndim(tensor)
    if len(factors) != nd:
        raise ValueError(f"got {len(factors)} factors for a tensor of order {nd}")
    rank = factors[0].shape[1] if _ndim(factors[0]) == 2 else 1
    for i, f in enumerate(factors):
        _lib.require_on_device(f, tensor.device, f"factors[{i}]")
        if i != mode and (f.shape[0] != tensor.shape[i] or (f.shape[1] if f.dim() == 2 else 1) != rank):
            raise ValueError(f"factor {i} has shape {tuple(f.shape)}, expected "
                             f"({tensor.shape[i]}, {rank})")
    if weights is not None:
        _lib.require_on_device(weights, tensor.device, "weights")
    fdt = _mttkrp_dtype(tensor, factors)
    x = tensor.contiguous()
    facs = [None if i == mode else f.reshape(f.shape[0], rank).to(fdt).contiguous()
            for i, f in enumerate(factors)]
    w = weights.to(fdt).contiguous() if weights is not None else None
    out = torch.empty((tensor.shape[mode], rank), dtype=fdt, device=tensor.device)
    xcode = _lib.dtype_code(x)
    prec = _lib.PREC_BF16 if xcode == _lib.BF16 else (_lib.PREC_VALU if xcode == _lib.F64 else _prec_code())
    sh, _ks = shape_array(x.shape)
    nbytes = lib().tla_mttkrp_workspace_size(xcode, nd, sh, rank, mode, prec)
    ws = workspace(nbytes, x.device)
    pp, _keep = ptr_array(facs)
    check(lib().tla_mttkrp(xcode, ptr(x), nd, sh, pp, ptr(w), rank, mode, prec, ptr(out), ptr(ws),
                           nbytes, stream(x.device)), "unfolding_dot_khatri_rao")
    return out


_COO_SEG = 1024  # nonzeros per piece of a row in the COO MTTKRP


def unfolding_dot_khatri_rao_coo(coords, values, shape, cp_tensor, mode):
    """MTTKRP of a sparse tensor given in coordinate form: ``coords`` (ndim x
    nnz int64 GPU tensor), ``values`` (nnz) and ``shape``.  Same result as
    :func:`unfolding_dot_khatri_rao` of the densified tensor (the path of
    ``tensorly/contrib/sparse/cp_tensor.py:9`` over the pydata ``sparse``
    backend), without densifying anything: the nonzeros are ordered by their
    ``mode`` index (a stable sort, so the result repeats bit for bit) and
    ``tla_mttkrp_coo`` sums each output row in that order."""
    weights, factors = cp_tensor
    _lib.require_gpu_tensor(values, "values")
    if coords.dim() != 2 or coords.shape[0] != len(shape) or coords.shape[1] != values.shape[0]:
        raise ValueError(f"coords must be (ndim, nnz) = ({len(shape)}, {values.shape[0]}), got {tuple(coords.shape)}")
    if len(factors) != len(shape):
        raise ValueError(f"{len(factors)} factors for an order-{len(shape)} tensor")
    dt = values.dtype
    if dt not in (torch.float32, torch.float64):
        raise TypeError(f"unfolding_dot_khatri_rao_coo takes float32 / float64 values, got {dt}")
    rank = int(factors[0].shape[1])
    coords = coords.to(device=values.device, dtype=torch.int64).contiguous()
    nnz = int(values.shape[0])
    for m, s in enumerate(shape):
        if m != mode:
            _lib.require_on_device(factors[m], values.device, f"factors[{m}]")
            if tuple(factors[m].shape) != (int(s), rank):
                raise ValueError(f"factor {m} has shape {tuple(factors[m].shape)}, expected ({int(s)}, {rank})")
    if weights is not None:
        _lib.require_on_device(weights, values.device, "weights")
    if nnz and (int(coords.min()) < 0 or any(int(coords[m].max()) >= int(s) for m, s in enumerate(shape))):
        raise ValueError("coordinates out of the tensor's shape")
    dev = values.device
    rows = int(shape[mode])
    idx = coords[mode]
    perm = torch.argsort(idx, stable=True) if nnz else torch.zeros(0, dtype=torch.int64, device=dev)
    counts = torch.bincount(idx, minlength=rows) if nnz else torch.zeros(rows, dtype=torch.int64, device=dev)
    row_ptr = torch.zeros(rows + 1, dtype=torch.int64, device=dev)
    row_ptr[1:] = torch.cumsum(counts, 0)
    # each row's nonzeros in pieces of at most _COO_SEG, so that modes with
    # few rows still spread over many waves
    nseg = (counts + (_COO_SEG - 1)) // _COO_SEG
    seg_row_ptr = torch.zeros(rows + 1, dtype=torch.int64, device=dev)
    seg_row_ptr[1:] = torch.cumsum(nseg, 0)
    n_seg = int(seg_row_ptr[-1])
    seg_row = torch.repeat_interleave(torch.arange(rows, device=dev), nseg)
    seg_start = row_ptr[seg_row] + (torch.arange(n_seg, device=dev) - seg_row_ptr[seg_row]) * _COO_SEG
    facs = [None if m == mode else factors[m].to(dt).contiguous() for m in range(len(shape))]
    w = None if weights is None else weights.to(dt).contiguous()
    out = torch.empty((rows, rank), dtype=dt, device=dev)
    code = _lib.dtype_code(values)
    nb = lib().tla_mttkrp_coo_workspace_size(code, n_seg, rank)
    ws = workspace(nb, dev)
    sh, _keep_sh = shape_array(shape)
    pp, _keep_f = ptr_array(facs)
    check(lib().tla_mttkrp_coo(code, nnz, len(shape), sh, ptr(coords), ptr(values.contiguous()), ptr(perm),
                               ptr(row_ptr), n_seg, ptr(seg_row), ptr(seg_start), ptr(seg_row_ptr), _COO_SEG, pp,
                               ptr(w), rank, mode, ptr(out), ptr(ws), nb, stream(dev)),
          "mttkrp_coo")
    return out


def mttkrp_partial(tensor, cp_tensor, keep, precision_name=None):
    """Dimension-tree partial MTTKRP: contract every mode outside the
    contiguous block ``keep = (begin, end)`` (a prefix or a suffix of the
    modes); returns a ``(prod kept sizes, R)`` matrix."""
    weights, factors = cp_tensor
    _lib.require_gpu_tensor(tensor, "tensor")
    nd = _ndim(tensor)
    kb, ke = keep
    for i, f in enumerate(factors):
        if not kb <= i < ke:
            _lib.require_on_device(f, tensor.device, f"factors[{i}]")
    if weights is not None:
        _lib.require_on_device(weights, tensor.device, "weights")
    rank = factors[0].shape[1]
    facs = [None if kb <= i < ke else f.to(torch.float32).contiguous() for i, f in enumerate(factors)]
    w = weights.to(torch.float32).contiguous() if weights is not None else None
    rows = 1
    for m in range(kb, ke):
        rows *= tensor.shape[m]
    out = torch.empty((rows, rank), dtype=torch.float32, device=tensor.device)
    xcode = _lib.dtype_code(tensor)
    prec = _lib.PREC_BF16 if xcode == _lib.BF16 else _prec_code(precision_name)
    sh, _ks = shape_array(tensor.shape)
    nbytes = lib().tla_mttkrp_partial_workspace_size(xcode, nd, sh, rank, kb, ke, prec)
    ws = workspace(nbytes, tensor.device)
    pp, _keep = ptr_array(facs)
    check(lib().tla_mttkrp_partial(xcode, ptr(tensor), nd, sh, pp, ptr(w), rank, kb, ke, prec, ptr(out),
                                   ptr(ws), nbytes, stream(tensor.device)), "mttkrp_partial")
    return out


def rankdiag_contract(Y, F, axis):
    """``Y`` is (I, J, R): axis 1 -> sum_j Y[i,j,r] F[j,r]; axis 0 -> sum_i Y[i,j,r] F[i,r]."""
    _lib.require_gpu_tensor(Y, "Y")
    _lib.require_on_device(F, Y.device, "F")
    I, J, R = Y.shape
    out = torch.empty((I if axis == 1 else J, R), dtype=torch.float32, device=Y.device)
    nbytes = lib().tla_rankdiag_contract_workspace_size(I, J, R, axis)
    ws = workspace(nbytes, Y.device)
    check(lib().tla_rankdiag_contract(ptr(Y), I, J, R, ptr(F), axis, ptr(out), ptr(ws), nbytes,
                                      stream(Y.device)), "rankdiag_contract")
    return out


# ---------------------------------------------------------------------------
def mode_dot(tensor, matrix_or_vector, mode, transpose=False):
    """n-mode product (``n_mode_product.py:5-76``): matrix -> mode size
    replaced by the matrix's row count (column count with ``transpose``);
    vector -> mode removed.  ValueError on mismatched sizes."""
    mv = matrix_or_vector
    new_shape = list(tensor.shape)
    if _ndim(mv) == 2:
        dim = 0 if transpose else 1
        if mv.shape[dim] != tensor.shape[mode]:
            raise ValueError(
                f"shapes {tuple(tensor.shape)} and {tuple(mv.shape)} not aligned in mode-{mode} multiplication: "
                f"{tensor.shape[mode]} (mode {mode}) != {mv.shape[dim]} (dim 1 of matrix)")
        J = mv.shape[1] if transpose else mv.shape[0]
        new_shape[mode] = J
        is_vec = False
    elif _ndim(mv) == 1:
        if mv.shape[0] != tensor.shape[mode]:
            raise ValueError(
                f"shapes {tuple(tensor.shape)} and {tuple(mv.shape)} not aligned for mode-{mode} multiplication: "
                f"{tensor.shape[mode]} (mode {mode}) != {mv.shape[0]} (vector size)")
        J = 1
        new_shape.pop(mode) if len(new_shape) > 1 else None
        if _ndim(tensor) == 1:
            new_shape = []
        is_vec = True
    else:
        raise ValueError("Can only take n_mode_product with a vector or a matrix."
                         f"Provided array of dimension {_ndim(mv)} not in [1, 2].")
    _lib.require_gpu_tensor(tensor, "tensor")
    _lib.require_on_device(mv, tensor.device, "matrix_or_vector")
    dt = torch.float64 if (tensor.dtype == torch.float64 or mv.dtype == torch.float64) else torch.float32
    x = tensor.to(dt).contiguous()
    u = mv.to(dt).contiguous()
    out_shape = list(tensor.shape)
    out_shape[mode] = J
    out = torch.empty(out_shape, dtype=dt, device=x.device)
    sh, _ks = shape_array(x.shape)
    prec = _lib.PREC_VALU if dt == torch.float64 else _prec_code()
    check(lib().tla_mode_dot(_lib.dtype_code(x), ptr(x), _ndim(x), sh, ptr(u), J, int(is_vec), mode,
                             int(bool(transpose)), prec, ptr(out), stream(x.device)), "mode_dot")
    return out.reshape(new_shape)


def multi_mode_dot(tensor, matrix_or_vec_list, modes=None, skip=None, transpose=False):
    """Chain of :func:`mode_dot` sorted by mode (``n_mode_product.py:79-135``);
    ``skip`` indexes the mode-sorted list; vector contractions shift later
    modes down."""
    if modes is None:
        modes = range(len(matrix_or_vec_list))
    pairs = sorted(zip(matrix_or_vec_list, modes), key=lambda p: p[1])
    res = tensor
    removed = 0
    for pos, (mv, mode) in enumerate(pairs):
        if skip is not None and pos == skip:
            continue
        res = mode_dot(res, mv, mode - removed, transpose=transpose)
        if _ndim(mv) == 1:
            removed += 1
    return res


# ---------------------------------------------------------------------------
def register(name="amd"):
    """Register these functions as a tensorly tenalg backend named ``name``
    (subclass of the core backend, so the seven out-of-path functions are
    inherited) and make it selectable with ``tensorly.tenalg.set_backend``.
    Returns the backend class, or None when tensorly is not importable."""
    try:
        import tensorly.tenalg as tlg
        from tensorly.tenalg.core_tenalg import CoreTenalgBackend
    except ImportError:
        return None

    class AmdTenalgBackend(CoreTenalgBackend, backend_name=name):
        pass

    for fn in (khatri_rao, unfolding_dot_khatri_rao, mode_dot, multi_mode_dot):
        AmdTenalgBackend.register_method(fn.__name__, fn)
    mgr = type(tlg)
    if name not in mgr.available_backend_names:
        mgr.available_backend_names.append(name)
    mgr._loaded_backends[name] = AmdTenalgBackend()
    return AmdTenalgBackend
