"""CP-ALS drop-in (``tensorly/decomposition/_cp.py``, 0.9.0) on the gfx950
kernels.

:func:`parafac` keeps the reference signature (``_cp.py:230-250``) and return
convention (``:532-543``); the mode loop runs in the native driver
(``tla_cp_als_*``): fused MTTKRP, FP64 Hadamard-of-Grams + SPD solve,
FP64 fit from the last mode's MTTKRP (``:217-225``), with no host
synchronisation unless ``tol``, ``callback``, ``verbose`` or the line search
ask for the error on the host (as the reference does).

Extensions (keyword-only, after the reference arguments):
``precision`` ("fp32" | "bf16" | "bf16x3"), ``dimtree`` (share partial
contractions between modes, orders 3 and 4) and ``extrapolation``
(``None``, or a :class:`tensorly_amd.learned.LearnedExtrapolation`).

``orthogonalise`` replaces each factor by the Q of its Householder QR
(``tla_qr``) before the sweeps it names (``_cp.py:330-331, 395-399``).

``mask`` keeps a working copy of the tensor whose missing entries are
re-imputed from the model after every sweep by one fused pass
(``tla_cp_impute``), which also gives the error (``_cp.py:194-207,
437-445``); ``init='svd'`` imputes the unfolding ``svd_mask_repeats`` times.

Not implemented (raise NotImplementedError rather than run elsewhere):
``sparsity``, complex tensors.
"""
import warnings

import numpy as np
import torch

from .. import _lib
from .._lib import check, lib, ptr, ptr_array, shape_array, stream
from ..cp_tensor import CPTensor, cp_normalize, grams64, validate_cp_rank
from ..tenalg import get_precision, unfolding_dot_khatri_rao

__all__ = ["parafac", "initialize_cp", "error_calc", "CP", "CpAlsEngine"]


def _check_random_state(seed):
    """``backend/core.py:162-188``."""
    if seed is None:
        return np.random.mtrand._rand
    if isinstance(seed, int):
        return np.random.RandomState(seed)
    if isinstance(seed, np.random.RandomState):
        return seed
    raise ValueError("Seed should be None, int or np.random.RandomState")


def initialize_cp(tensor, rank, init="svd", svd="truncated_svd", non_negative=False, random_state=None,
                  normalize_factors=False, mask=None, svd_mask_repeats=5):
    """``_cp.py:26-137``.  ``init='random'`` draws ``random_sample((I_m, R))``
    per mode in order from the RandomState (``random/base.py:103-105``) and
    casts to the tensor's dtype on its device; ``init='svd'`` uses the GPU
    Gram eigensolver (:func:`tensorly_amd.svd.svd_interface`), with the
    mask's missing entries imputed ``svd_mask_repeats`` times when a mask is
    given (``:66-76``)."""
    rng = _check_random_state(random_state)
    dt, dev = tensor.dtype, tensor.device
    if isinstance(init, str) and init == "random":
        factors = [torch.tensor(rng.random_sample((s, rank)), dtype=dt, device=dev) for s in tensor.shape]
        weights = torch.ones(rank, dtype=dt, device=dev)
    elif isinstance(init, str) and init == "svd":
        from ..svd import svd_interface
        from ..base import unfold
        factors = []
        for mode in range(tensor.dim()):
            U, S, _ = svd_interface(unfold(tensor, mode), n_eigenvecs=rank, method=svd, non_negative=non_negative,
                                    mask=None if mask is None else unfold(mask, mode),
                                    n_iter_mask_imputation=svd_mask_repeats)
            if mode == 0:
                k = min(rank, S.shape[0])
                U = U.clone()
                U[:, :k] = U[:, :k] * S[:k]
            if tensor.shape[mode] < rank:
                pad = torch.tensor(rng.random_sample((U.shape[0], rank - tensor.shape[mode])), dtype=dt,
                                   device=dev)
                U = torch.cat([U, pad], dim=1)
            factors.append(U[:, :rank].contiguous())
        weights = torch.ones(rank, dtype=dt, device=dev)
    elif isinstance(init, (tuple, list, CPTensor)):
        try:
            if normalize_factors is True:
                warnings.warn("It is not recommended to initialize a tensor with normalizing. Consider "
                              "normalizing the tensor before using this function")
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
    if non_negative:
        factors = [torch.abs(f) for f in factors]
    kt = CPTensor((weights, factors))
    if normalize_factors:
        kt = cp_normalize(kt)
    return kt


def _inner_with_last_mttkrp(tensor, weights, factors):
    """``<X, [|w; A|]>`` = sum(MTTKRP_last(A, w) * A_last): one fused pass."""
    m = unfolding_dot_khatri_rao(tensor, (weights, factors), tensor.dim() - 1)
    return torch.sum(m.to(torch.float64) * factors[-1].to(torch.float64))


def _cp_norm2_64(weights, factors):
    gs = grams64(factors)
    acc = gs[0]
    for g in gs[1:]:
        acc = acc * g
    w = weights.to(torch.float64)
    return torch.sum(acc * (w[:, None] * w[None, :]))


def error_calc(tensor, norm_tensor, weights, factors, sparsity=None, mask=None, mttkrp=None):
    """``_cp.py:162-227`` (no sparsity).  Both branches use
    ``sqrt(|‖X‖² + ‖M‖² − 2<X, M>|)``; without ``mttkrp`` the inner product
    comes from one fused MTTKRP pass instead of a full reconstruction (same
    quantity).  With ``mask`` the tensor's missing entries are replaced by
    the model first (a new tensor) and its norm recomputed, in one
    ``tla_cp_impute`` pass.  Returns ``(unnormalised error, tensor,
    norm_tensor)``."""
    if sparsity:
        raise NotImplementedError("sparsity is not supported by tensorly_amd")
    if mask is not None:
        _lib.require_on_device(mask, tensor.device, "mask")
        work = tensor.contiguous().clone()
        err, nrm = _Imputer(work, mask.to(tensor.dtype).contiguous()).run(weights, factors, write=True)
        return err, work, nrm
    nm2 = _cp_norm2_64(weights, factors)
    if mttkrp is None:
        ip = _inner_with_last_mttkrp(tensor, weights, factors)
    else:
        ip = torch.sum(mttkrp.to(torch.float64) * factors[-1].to(torch.float64))
    nx = torch.as_tensor(norm_tensor, dtype=torch.float64, device=tensor.device)
    err = torch.sqrt(torch.abs(nx * nx + nm2 - 2 * ip))
    return err, tensor, norm_tensor


class CpAlsEngine:
    """Owns the native sweep driver for one tensor and a set of factor
    buffers (updated in place)."""

    def __init__(self, tensor, weights, factors, l2_reg=0.0, precision=None, dimtree=False,
                 fixed_modes=None, tensor_bf16=None, unit_weights=False):
        """``unit_weights``: the weights are ones and stay ones (no
        normalisation), so the native driver skips every weight product."""
        self.tensor = tensor
        self.dev = tensor.device
        self.nd = tensor.dim()
        self.rank = factors[0].shape[1]
        self.factors = [f.contiguous().clone() for f in factors]
        self.weights = weights.contiguous().clone()
        prec = precision or get_precision()
        self.precision = prec
        x = tensor if tensor_bf16 is None else tensor_bf16
        self.x = x.contiguous()
        xcode = _lib.dtype_code(self.x)
        pcode = _lib.PRECISIONS[prec]
        if xcode == _lib.F64:
            pcode = _lib.PREC_VALU
        if xcode == _lib.BF16:
            pcode = _lib.PREC_BF16
        self._sh, self._keep_sh = shape_array(self.x.shape)
        self.dimtree = bool(dimtree) and bool(lib().tla_cp_als_dimtree_supported(
            xcode, self.nd, self._sh, self.rank, pcode))
        nb = lib().tla_cp_als_workspace_size(xcode, self.nd, self._sh, self.rank, pcode, int(self.dimtree))
        if nb == 0:
            raise ValueError(f"unsupported CP-ALS configuration (order {self.nd}, rank {self.rank}; "
                             "rank must be in [1, 64])")
        self.ws = torch.empty(nb, dtype=torch.uint8, device=self.dev)
        fixed = (_lib.c_i * self.nd)(*[int(m in (fixed_modes or [])) for m in range(self.nd)])
        self._fixed = fixed
        fp, self._keep_f = ptr_array(self.factors)
        handle = _lib.c_p()
        self.unit_weights = bool(unit_weights)
        wptr = _lib.c_p(0) if self.unit_weights else ptr(self.weights)
        check(lib().tla_cp_als_create(xcode, ptr(self.x), self.nd, self._sh, self.rank, fp, wptr,
                                      float(l2_reg), pcode, int(self.dimtree), fixed, ptr(self.ws), nb,
                                      _lib.ctypes.byref(handle)), "cp_als_create")
        self.handle = handle
        if xcode == _lib.BF16:
            # ‖X‖² of the bf16 copy: the run solves the rounded tensor's
            # problem, and its errors are that problem's (the float32
            # original's differ by about the rounding's ‖X - X_bf16‖)
            nx2 = torch.sum(self.x.to(torch.float64) ** 2).reshape(1)
            self.refresh(with_norm=False)
            check(lib().tla_cp_als_set_normx2(self.handle, ptr(nx2), stream(self.dev)), "cp_als_set_normx2")
        else:
            self.refresh(with_norm=True)

    def refresh(self, with_norm=False):
        check(lib().tla_cp_als_refresh(self.handle, int(with_norm), stream(self.dev)), "cp_als_refresh")

    def sweeps(self, n, errors=None, first_slot=0, normalize=False):
        check(lib().tla_cp_als_sweeps(self.handle, int(n), ptr(errors), int(first_slot), int(bool(normalize)),
                                      stream(self.dev)), "cp_als_sweeps")

    def normx2(self):
        """Device FP64 ``‖X‖²`` (1-element tensor, no host sync)."""
        buf = torch.empty(1, dtype=torch.float64, device=self.dev)
        check(lib().tla_cp_als_copy_state(self.handle, 1, ptr(buf), 8, stream(self.dev)), "cp_als_copy_state")
        return buf[0]

    def raise_if_singular(self):
        """Raise ``numpy.linalg.LinAlgError`` if a solve since the last
        refresh met a singular normal-equations matrix: the error that
        ``tl.solve`` (``np.linalg.solve``) raises in the reference's mode
        update (``_cp.py:425-427``).  Synchronises with the device."""
        buf = torch.zeros(1, dtype=torch.int32, device=self.dev)
        check(lib().tla_cp_als_copy_state(self.handle, 4, ptr(buf), 4, stream(self.dev)), "cp_als_copy_state")
        if int(buf) == 2:
            raise np.linalg.LinAlgError("Singular matrix")

    def set_factors(self, weights, factors):
        self.weights.copy_(weights)
        for f, g in zip(self.factors, factors):
            f.copy_(g)
        self.refresh(with_norm=False)

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().tla_cp_als_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Imputer:
    """The working tensor of ``parafac(..., mask=...)``: after each sweep its
    missing entries are replaced by the model (``_cp.py:194-207, 437-445``)
    by one ``tla_cp_impute`` pass, which also gives the imputed tensor's
    norm and the reconstruction error."""

    def __init__(self, work, mask):
        self.work, self.mask, self.dev = work, mask, work.device
        self._sh, self._keep_sh = shape_array(work.shape)
        self.nb = lib().tla_cp_impute_workspace_size(_lib.dtype_code(work), work.dim(), self._sh)
        self.ws = _lib.workspace(self.nb, self.dev)
        self.sums = torch.zeros(2, dtype=torch.float64, device=self.dev)

    def run(self, weights, factors, write):
        """Impute from the model ``(weights, factors)`` (``write``: into the
        working tensor); returns device ``(error, norm)`` of the imputed
        tensor."""
        facs = [f.to(self.work.dtype).contiguous() for f in factors]
        w = weights.to(self.work.dtype).contiguous()
        pp, _keep = ptr_array(facs)
        check(lib().tla_cp_impute(_lib.dtype_code(self.work), ptr(self.work), ptr(self.mask), self.work.dim(),
                                  self._sh, pp, ptr(w), facs[0].shape[1], int(bool(write)), ptr(self.sums),
                                  ptr(self.ws), self.nb, stream(self.dev)), "cp_impute")
        s = torch.sqrt(self.sums)
        return s[1], s[0]


def parafac(tensor, rank, n_iter_max=100, init="svd", svd="truncated_svd", normalize_factors=False,
            orthogonalise=False, tol=1e-8, random_state=None, verbose=0, return_errors=False,
            sparsity=None, l2_reg=0, mask=None, cvg_criterion="abs_rec_error", fixed_modes=None,
            svd_mask_repeats=5, linesearch=False, callback=None, *, precision=None, dimtree=False,
            extrapolation=None, tensor_bf16=None):
    """CANDECOMP/PARAFAC by alternating least squares (``_cp.py:230-543``).

    Returns ``CPTensor((weights, factors))`` and, with ``return_errors``, the
    list of relative reconstruction errors (Python floats)."""
    if sparsity:
        raise NotImplementedError("sparsity is not supported by tensorly_amd.parafac")
    if orthogonalise and not isinstance(orthogonalise, int):
        orthogonalise = n_iter_max  # (True is an int: iterations 0 and 1, as in the reference)
    _lib.require_gpu_tensor(tensor, "tensor")
    if tensor.dtype not in (torch.float32, torch.float64):
        raise TypeError("parafac needs a float32 or float64 tensor")
    rank = validate_cp_rank(tuple(tensor.shape), rank=rank)
    if linesearch and extrapolation is not None:
        raise ValueError("choose either the reference line search or a learned extrapolation")
    if orthogonalise and extrapolation is not None:
        raise NotImplementedError("orthogonalise runs the reference's host loop, not the extrapolated driver")

    imputer = None
    if mask is not None:
        _lib.require_on_device(mask, tensor.device, "mask")
        if tuple(mask.shape) != tuple(tensor.shape):
            raise ValueError(f"mask has shape {tuple(mask.shape)}, the tensor {tuple(tensor.shape)}")
        if extrapolation is not None or tensor_bf16 is not None:
            raise NotImplementedError("mask runs the reference's host loop on a float32 / float64 working copy")
        mask = mask.to(tensor.dtype).contiguous()
    weights, factors = initialize_cp(tensor, rank, init=init, svd=svd, random_state=random_state,
                                     normalize_factors=normalize_factors, mask=mask,
                                     svd_mask_repeats=svd_mask_repeats)
    if mask is not None:
        tensor = tensor.contiguous().clone()  # the working copy whose missing entries are imputed
        imputer = _Imputer(tensor, mask)
    if weights is None:
        weights = torch.ones(rank, dtype=tensor.dtype, device=tensor.device)
    nd = tensor.dim()
    fixed_modes = list(fixed_modes) if fixed_modes is not None else []
    if fixed_modes == list(range(nd)):
        return CPTensor((weights, factors))
    if nd - 1 in fixed_modes:
        warnings.warn("You asked for fixing the last mode, which is not supported.\n The last mode will "
                      "not be fixed. Consider using tl.moveaxis()")
        fixed_modes.remove(nd - 1)

    unit = (not normalize_factors) and bool(torch.all(weights == 1))
    eng = CpAlsEngine(tensor, weights, factors, l2_reg=l2_reg, precision=precision, dimtree=dimtree,
                      fixed_modes=fixed_modes, tensor_bf16=tensor_bf16, unit_weights=unit)
    try:
        if extrapolation is not None:
            from ..learned import run_extrapolated_als
            return run_extrapolated_als(eng, n_iter_max, tol, extrapolation, return_errors,
                                        normalize_factors, cvg_criterion, callback, verbose)
        return _run_plain(eng, tensor, n_iter_max, tol, return_errors, normalize_factors, cvg_criterion,
                          linesearch, callback, verbose, orthogonalise=orthogonalise, imputer=imputer)
    finally:
        eng.close()


def _orthogonalise(eng):
    """``_cp.py:395-399``: every factor with at least ``rank`` rows replaced
    by the Q of its reduced QR (Householder, LAPACK's signs)."""
    from ..svd import qr
    rank = eng.rank
    eng.set_factors(eng.weights, [qr(f)[0] if min(f.shape) >= rank else f for f in eng.factors])


def _run_plain(eng, tensor, n_iter_max, tol, return_errors, normalize, cvg_criterion, linesearch, callback,
               verbose, orthogonalise=False, imputer=None):
    dev = tensor.device
    errs_dev = torch.zeros(max(1, n_iter_max), dtype=torch.float64, device=dev)
    want_err = bool(tol or return_errors)
    # the device-side sweep loop normalises float32 factors only
    host_loop = (bool(tol) or callback is not None or linesearch or verbose or bool(orthogonalise)
                 or imputer is not None or (normalize and tensor.dtype != torch.float32))
    # only the stopping rule needs the host: check it once per chunk of
    # sweeps and roll back to the stopping sweep's factors
    chunked = (bool(tol) and callback is None and not linesearch and not verbose and not orthogonalise
               and imputer is None and not normalize)
    if chunked:
        return _run_chunked(eng, n_iter_max, tol, return_errors, cvg_criterion, errs_dev)
    norm_tensor = torch.sqrt(eng.normx2())

    if callback is not None:
        if imputer is not None:
            e0, norm_tensor = imputer.run(eng.weights, eng.factors, write=False)
        else:
            e0, _, _ = error_calc(tensor, norm_tensor, eng.weights, eng.factors)
        callback(CPTensor((eng.weights.clone(), [f.clone() for f in eng.factors])), float(e0 / norm_tensor))

    if not host_loop:
        if n_iter_max > 0:
            eng.sweeps(n_iter_max, errs_dev if want_err else None, 0, normalize)
        eng.raise_if_singular()
        rec_errors = [float(v) for v in errs_dev[:n_iter_max].cpu()] if want_err else []
        return _result(eng, rec_errors, return_errors)

    rec_errors = []
    if linesearch:
        acc_pow, acc_fail, max_fail = 2.0, 0, 4
    rec_error = None
    for iteration in range(n_iter_max):
        if orthogonalise and iteration <= orthogonalise:
            _orthogonalise(eng)
        if linesearch and iteration % 2 == 0:
            factors_last = [f.clone() for f in eng.factors]
            weights_last = eng.weights.clone()
        eng.sweeps(1, errs_dev if imputer is None else None, iteration, False)
        eng.raise_if_singular()
        line_iter = bool(linesearch and iteration % 2 == 0 and iteration > 5)
        if imputer is not None:
            # the working tensor is re-imputed from the updated model whether
            # or not the error is wanted (_cp.py:437-445)
            unnorm, norm_i = imputer.run(eng.weights, eng.factors, write=True)
            if want_err and not line_iter:
                norm_tensor = norm_i
        if line_iter:
            jump = iteration ** (1.0 / acc_pow)
            new_w = weights_last + (eng.weights - weights_last) * jump
            new_f = [fl + (f - fl) * jump for f, fl in zip(eng.factors, factors_last)]
            if imputer is not None:
                new_err, new_norm = imputer.run(new_w, new_f, write=False)
            else:
                new_err, _, _ = error_calc(tensor, norm_tensor, new_w, new_f)
                new_norm = norm_tensor
            if float(new_err / new_norm) < rec_errors[-1]:
                eng.set_factors(new_w, new_f)
                if imputer is not None:
                    imputer.run(new_w, new_f, write=True)
                    norm_tensor = new_norm
                acc_fail = 0
                if verbose:
                    print(f"Accepted line search jump of {jump}.")
            else:
                if imputer is not None:
                    _, norm_tensor = imputer.run(eng.weights, eng.factors, write=True)
                acc_fail += 1
                if verbose:
                    print(f"Line search failed for jump of {jump}.")
                if acc_fail == max_fail:
                    acc_pow += 1.0
                    acc_fail = 0
        elif want_err:
            rec_error = float(unnorm / norm_tensor) if imputer is not None else float(errs_dev[iteration])
            rec_errors.append(rec_error)
        if callback is not None:
            cp = CPTensor((eng.weights.clone(), [f.clone() for f in eng.factors]))
            if callback(cp, rec_error) is True:
                break
        if tol and iteration >= 1:
            dec = rec_errors[-2] - rec_errors[-1]
            if verbose:
                print(f"iteration {iteration}, reconstruction error: {rec_error}, decrease = {dec}")
            if cvg_criterion == "abs_rec_error":
                stop = abs(dec) < tol
            elif cvg_criterion == "rec_error":
                stop = dec < tol
            else:
                raise TypeError("Unknown convergence criterion")
            if stop:
                if verbose:
                    print(f"PARAFAC converged after {iteration} iterations")
                break
        if normalize:
            w, f = cp_normalize((eng.weights, eng.factors))
            eng.set_factors(w, f)
    return _result(eng, rec_errors, return_errors)


# sweeps per host check of the stopping rule
_TOL_CHUNK = 8


def _run_chunked(eng, n_iter_max, tol, return_errors, cvg_criterion, errs_dev):
    """The plain ALS loop with ``tol`` (``_cp.py:484-514``) with one host
    synchronisation per ``_TOL_CHUNK`` sweeps instead of one per sweep: the
    sweeps of a chunk run back to back on the device, each followed by a copy
    of the factors; the stopping rule is then applied sweep by sweep on the
    recorded errors, and at the stopping sweep the factors are restored from
    its copy, so the result is the one the per-sweep check gives."""
    if cvg_criterion not in ("abs_rec_error", "rec_error"):
        raise TypeError("Unknown convergence criterion")
    rec_errors = []
    it = 0
    while it < n_iter_max:
        k = min(_TOL_CHUNK, n_iter_max - it)
        snaps = []
        for j in range(k):
            eng.sweeps(1, errs_dev, it + j, False)
            snaps.append([f.clone() for f in eng.factors])
        errs = [float(v) for v in errs_dev[it:it + k].cpu()]
        eng.raise_if_singular()
        for j in range(k):
            rec_errors.append(errs[j])
            i = it + j
            if i >= 1:
                dec = rec_errors[-2] - rec_errors[-1]
                stop = abs(dec) < tol if cvg_criterion == "abs_rec_error" else dec < tol
                if stop:
                    if j < k - 1:
                        eng.set_factors(eng.weights, snaps[j])
                    return _result(eng, rec_errors, return_errors)
        it += k
    return _result(eng, rec_errors, return_errors)


def _result(eng, rec_errors, return_errors):
    cp = CPTensor((eng.weights.clone(), [f.clone() for f in eng.factors]))
    if return_errors:
        return cp, rec_errors
    return cp


class CP:
    """Scikit-style wrapper (``_cp.py:765-920``): ``CP(rank, ...).fit_transform(X)``."""

    def __init__(self, rank, n_iter_max=100, init="svd", svd="truncated_svd", normalize_factors=False,
                 orthogonalise=False, tol=1e-8, random_state=None, verbose=0, sparsity=None, l2_reg=0,
                 mask=None, cvg_criterion="abs_rec_error", fixed_modes=None, svd_mask_repeats=5,
                 linesearch=False, callback=None):
        self.rank, self.n_iter_max, self.init, self.svd = rank, n_iter_max, init, svd
        self.normalize_factors, self.orthogonalise, self.tol = normalize_factors, orthogonalise, tol
        self.random_state, self.verbose, self.sparsity, self.l2_reg = random_state, verbose, sparsity, l2_reg
        self.mask, self.cvg_criterion, self.fixed_modes = mask, cvg_criterion, fixed_modes
        self.svd_mask_repeats, self.linesearch, self.callback = svd_mask_repeats, linesearch, callback

    def fit_transform(self, tensor):
        cp, errors = parafac(tensor, rank=self.rank, n_iter_max=self.n_iter_max, init=self.init, svd=self.svd,
                             normalize_factors=self.normalize_factors, orthogonalise=self.orthogonalise,
                             tol=self.tol, random_state=self.random_state, verbose=self.verbose,
                             sparsity=self.sparsity, l2_reg=self.l2_reg, mask=self.mask,
                             cvg_criterion=self.cvg_criterion, fixed_modes=self.fixed_modes,
                             svd_mask_repeats=self.svd_mask_repeats, linesearch=self.linesearch,
                             callback=self.callback, return_errors=True)
        self.decomposition_ = cp
        self.errors_ = errors
        return self.decomposition_

    def fit(self, tensor):
        self.fit_transform(tensor)
        return self

    def __repr__(self):
        return f"Rank-{self.rank} CP decomposition."
