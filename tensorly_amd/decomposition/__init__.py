"""Decomposition drop-ins: :func:`parafac` / :class:`CP` (CP-ALS),
:func:`non_negative_parafac` (multiplicative updates),
:func:`non_negative_parafac_hals` (HALS), :func:`constrained_parafac`
(AO-ADMM) and :func:`tucker` /
:func:`partial_tucker` / :class:`Tucker` (HOOI)."""
from ._constrained_cp import ConstrainedCP, constrained_parafac, initialize_constrained_parafac  # noqa: F401
from ._cp import CP, error_calc, initialize_cp, parafac  # noqa: F401
from ._nn_cp import non_negative_parafac, non_negative_parafac_hals  # noqa: F401
from ._sparse_cp import parafac_coo  # noqa: F401
from ._tucker import Tucker, initialize_tucker, partial_tucker, tucker  # noqa: F401
