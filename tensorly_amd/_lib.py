"""ctypes binding of ``libtensorly_amd.so`` (the C ABI declared in
``include/tensorly_amd.h``).

torch is imported first on purpose: the library's NEEDED entry for the HIP
runtime (``libamdhip64.so.7``) then resolves to the copy torch already loaded,
so both share one runtime, one device context and torch's streams.

There is no CPU fallback: if the library is missing or no GPU is present,
every compute entry point raises.
"""
import ctypes
import os
import threading

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtensorly_amd.so")

# status codes / enums (mirror include/tensorly_amd.h)
OK, ERR_ARG, ERR_HIP, ERR_UNSUPPORTED, ERR_NUMERIC = 0, 1, 2, 3, 4
F32, F64, BF16 = 0, 1, 2
PREC_FP32, PREC_BF16, PREC_BF16X3, PREC_VALU = 0, 1, 2, 3
PRECISIONS = {"fp32": PREC_FP32, "bf16": PREC_BF16, "bf16x3": PREC_BF16X3, "valu": PREC_VALU}

_lib = None
_lock = threading.Lock()

c_p = ctypes.c_void_p
c_i = ctypes.c_int
c_i64 = ctypes.c_int64
c_sz = ctypes.c_size_t
c_d = ctypes.c_double
c_f = ctypes.c_float
P_i64 = ctypes.POINTER(ctypes.c_int64)
PP = ctypes.POINTER(ctypes.c_void_p)

_SIGNATURES = {
    "tla_last_error": (ctypes.c_char_p, []),
    "tla_abi_version": (c_i, []),
    "tla_mttkrp_workspace_size": (c_sz, [c_i, c_i, P_i64, c_i, c_i, c_i]),
    "tla_mttkrp": (c_i, [c_i, c_p, c_i, P_i64, PP, c_p, c_i, c_i, c_i, c_p, c_p, c_sz, c_p]),
    "tla_mttkrp_noalloc_max_scratch": (c_sz, []),
    "tla_mttkrp_noalloc": (c_i, [c_i, c_p, c_i, P_i64, PP, c_p, c_i, c_i, c_i, c_p, c_p]),
    "tla_mttkrp_partial_workspace_size": (c_sz, [c_i, c_i, P_i64, c_i, c_i, c_i, c_i]),
    "tla_mttkrp_partial_supported": (c_i, [c_i, c_i, P_i64, c_i, c_i, c_i, c_i]),
    "tla_mttkrp_partial": (c_i, [c_i, c_p, c_i, P_i64, PP, c_p, c_i, c_i, c_i, c_i, c_p, c_p, c_sz, c_p]),
    "tla_rankdiag_contract_workspace_size": (c_sz, [c_i64, c_i64, c_i, c_i]),
    "tla_rankdiag_contract": (c_i, [c_p, c_i64, c_i64, c_i, c_p, c_i, c_p, c_p, c_sz, c_p]),
    "tla_khatri_rao": (c_i, [c_i, PP, c_i, P_i64, c_i, c_p, c_p, c_p, c_p]),
    "tla_mode_dot": (c_i, [c_i, c_p, c_i, P_i64, c_p, c_i64, c_i, c_i, c_i, c_i, c_p, c_p]),
    "tla_gram64_workspace_size": (c_sz, [c_i64, c_i]),
    "tla_gram64": (c_i, [c_i, c_p, c_i64, c_i, c_p, c_p, c_sz, c_p]),
    "tla_cp_solve_workspace_size": (c_sz, [c_i64, c_i]),
    "tla_cp_solve": (c_i, [c_i, PP, c_i, c_i, c_p, c_d, c_p, c_i64, c_i, c_p, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    "tla_sqnorm_workspace_size": (c_sz, [c_i64]),
    "tla_sqnorm": (c_i, [c_i, c_p, c_i64, c_p, c_p, c_sz, c_p]),
    "tla_cp_error": (c_i, [PP, c_i, c_p, c_i, c_i, c_p, c_p, c_p, c_i, c_p]),
    "tla_nn_mu_update": (c_i, [c_i, c_p, c_p, c_i64, c_i, PP, c_i, c_i, c_p, c_i, c_d, c_p]),
    "tla_cp_lstsq_grad_mode": (c_i, [c_i, c_p, c_p, c_p, c_i64, c_i, PP, c_i, c_i, c_p, c_i, c_p]),
    "tla_hals_nnls_workspace_size": (c_sz, [c_i64, c_i]),
    "tla_hals_nnls": (c_i, [c_i, c_p, c_p, c_i64, c_i, PP, c_i, c_i, c_p, c_i, c_d, c_i, c_i, c_d, c_p, c_p, c_sz,
                            c_p]),
    "tla_mttkrp_coo_workspace_size": (c_sz, [c_i, c_i64, c_i]),
    "tla_mttkrp_coo": (c_i, [c_i, c_i64, c_i, P_i64, c_p, c_p, c_p, c_p, c_i64, c_p, c_p, c_p, c_i64, PP, c_p, c_i,
                             c_i, c_p, c_p, c_sz, c_p]),
    "tla_cp_impute_workspace_size": (c_sz, [c_i, c_i, P_i64]),
    "tla_cp_impute": (c_i, [c_i, c_p, c_p, c_i, P_i64, PP, c_p, c_i, c_i, c_p, c_p, c_sz, c_p]),
    "tla_masked_blend": (c_i, [c_i, c_p, c_p, c_p, c_p, c_i64, c_p]),
    "tla_qr_workspace_size": (c_sz, [c_i64, c_i]),
    "tla_qr": (c_i, [c_i, c_p, c_i64, c_i, c_p, c_p, c_p, c_sz, c_p]),
    "tla_prox_workspace_size": (c_sz, [c_i64, c_i]),
    "tla_prox": (c_i, [c_i, c_p, c_p, c_i64, c_i, c_i, c_d, c_p, c_sz, c_p]),
    "tla_admm_workspace_size": (c_sz, [c_i64, c_i]),
    "tla_admm_update": (c_i, [c_i, c_p, c_p, c_p, c_p, c_i64, c_i, c_p, PP, c_i, c_i, c_i, c_d, c_i, c_d, c_p, c_p, c_p,
                              c_p, c_p, c_sz, c_p]),
    "tla_mode_gram64_workspace_size": (c_sz, [c_i, P_i64, c_i]),
    "tla_mode_gram64": (c_i, [c_i, c_p, c_i, P_i64, c_i, c_p, c_p, c_sz, c_p]),
    "tla_sym_eig_workspace_size": (c_sz, [c_i]),
    "tla_sym_eig": (c_i, [c_p, c_i, c_i, c_p, c_i, c_p, c_p, c_d, c_i, c_p, c_p, c_sz, c_p]),
    "tla_sym_eig_batch": (c_i, [c_i, PP, c_p, c_p, PP, c_p, PP, PP, c_d, c_i, PP, c_p, c_p]),
    "tla_gram_svd_block": (c_i, [c_i, c_i]),
    "tla_gram_left_singular": (c_i, [c_p, c_i, c_i, c_p, c_i, c_i, c_p, c_p, c_p, c_p, c_d, c_i, c_p, c_p, c_sz,
                                     c_p]),
    "tla_cp_als_workspace_size": (c_sz, [c_i, c_i, P_i64, c_i, c_i, c_i]),
    "tla_cp_als_dimtree_supported": (c_i, [c_i, c_i, P_i64, c_i, c_i]),
    "tla_cp_als_create": (c_i, [c_i, c_p, c_i, P_i64, c_i, PP, c_p, c_d, c_i, c_i, ctypes.POINTER(c_i), c_p,
                                c_sz, ctypes.POINTER(c_p)]),
    "tla_cp_als_dimtree_active": (c_i, [c_p]),
    "tla_cp_als_refresh": (c_i, [c_p, c_i, c_p]),
    "tla_cp_als_set_normx2": (c_i, [c_p, c_p, c_p]),
    "tla_cp_als_sweeps": (c_i, [c_p, c_i, c_p, c_i, c_i, c_p]),
    "tla_cp_als_copy_state": (c_i, [c_p, c_i, c_p, c_sz, c_p]),
    "tla_cp_als_destroy": (None, [c_p]),
    "tla_mlp_num_params": (c_i, []),
    "tla_mlp_forward": (c_i, [c_p, c_p, c_i, c_p, c_p]),
    "tla_mlp_train": (c_i, [c_p, c_p, c_p, c_p, c_p, c_p, c_i, c_i, c_i, c_f, c_f, c_f, c_f, c_i, c_p, c_p, c_p]),
    "tla_cp_als_extrap_workspace_size": (c_sz, [c_p]),
    "tla_cp_als_run_extrapolated": (c_i, [c_p, c_i, c_p, c_f, c_f, c_i, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    "tla_extrap_features": (c_i, [c_p, c_p, c_p, c_i, c_p, c_p]),
    "tla_batched_cp_max_blocks_per_cu": (c_i, [c_i]),
    "tla_batched_cp_als": (c_i, [c_i, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_i, c_i, c_d, c_i, c_p, c_f,
                                 c_f, c_p, c_i, c_i, c_i, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_i,
                                 c_i, c_i, c_p]),
}


def exported_symbols():
    """Names the header declares and this binding resolves."""
    return sorted(_SIGNATURES)


def lib():
    """Load (once) and return the native library; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"tensorly_amd native library not built ({LIB_PATH}); "
                    "run `python -m tensorly_amd.build`")
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def check(rc, what=""):
    if rc == OK:
        return
    msg = lib().tla_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == ERR_ARG:
        raise ValueError(text)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    raise RuntimeError(text)


def require_gpu_tensor(t, name="tensor"):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor on a GPU, got {type(t).__name__}")
    if t.device.type != "cuda":
        raise TypeError(f"{name} must live on a GPU (got device {t.device}); "
                        "tensorly_amd has no CPU path")
    if t.is_complex():
        raise NotImplementedError("tensorly_amd kernels are real-valued only")


def require_on_device(t, device, name):
    """``t`` must be a GPU tensor on ``device`` (the device of the call's
    main operand): a pointer from another device or from host memory must
    never reach a kernel."""
    require_gpu_tensor(t, name)
    if t.device != device:
        raise TypeError(f"{name} is on {t.device}, but the operation runs on {device}")


def dtype_code(t):
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.float64:
        return F64
    if t.dtype == torch.bfloat16:
        return BF16
    raise TypeError(f"unsupported dtype {t.dtype}; use float32 or float64")


def ptr(t):
    return c_p(t.data_ptr()) if t is not None else c_p(0)


def ptr_array(ts):
    arr = (c_p * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = t.data_ptr() if t is not None else 0
    return ctypes.cast(arr, PP), arr


def shape_array(shape):
    arr = (ctypes.c_int64 * len(shape))(*[int(s) for s in shape])
    return ctypes.cast(arr, P_i64), arr


def stream(device=None):
    return c_p(torch.cuda.current_stream(device).cuda_stream)


def workspace(nbytes, device):
    if nbytes == 0:
        return None
    return torch.empty(int(nbytes), dtype=torch.uint8, device=device)
