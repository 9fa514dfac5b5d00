"""tensorly_amd -- MI355X (gfx950) implementation of TensorLy's CP-ALS and
Tucker-HOOI hot path.

Drop-ins keep the reference signatures (tensorly 0.9.0):

* tenalg backend functions: :func:`khatri_rao`, :func:`unfolding_dot_khatri_rao`,
  :func:`mode_dot`, :func:`multi_mode_dot` (:func:`tenalg.register` plugs them
  into tensorly's tenalg dispatcher when tensorly is installed);
* decompositions: :func:`parafac`, :func:`non_negative_parafac`,
  :func:`non_negative_parafac_hals`, :func:`constrained_parafac`,
  :func:`tucker`, :func:`partial_tucker`.

Tensors are torch tensors on the GPU; the arithmetic runs in hand-written HIP
kernels reached through the C ABI of ``libtensorly_amd.so``
(``include/tensorly_amd.h``).  There is no CPU fallback.
"""
__version__ = "0.1.0"

from . import _lib  # noqa: F401  (imports torch before the HIP library)
from .base import fold, unfold  # noqa: F401
from .cp_tensor import CPTensor, cp_lstsq_grad, cp_norm, cp_normalize, cp_to_tensor, validate_cp_rank  # noqa: F401
from .tenalg import (khatri_rao, mode_dot, multi_mode_dot, precision, set_precision,  # noqa: F401
                     unfolding_dot_khatri_rao)
from .tucker_tensor import TuckerTensor, tucker_to_tensor, validate_tucker_rank  # noqa: F401
from .decomposition import (CP, ConstrainedCP, Tucker, constrained_parafac, non_negative_parafac,  # noqa: F401
                            non_negative_parafac_hals, parafac, parafac_coo, partial_tucker, tucker)


def native_library_path():
    return _lib.LIB_PATH
