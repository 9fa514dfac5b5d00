"""CPU FP64 oracle for the CP-ALS / Tucker-HOOI path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
directory; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker (or, for
``bench.py``, as the CPU baseline being timed).

``oracle.tensorly_np`` restates, in plain numpy, the reference TensorLy 0.9.0
numerics for the functions on the path (each function cites the reference
file:line it follows).  ``oracle.pinning`` lists how each function is pinned
(golden vectors from the reference's own tests, closed forms, brute force,
LAPACK identities, or the committed trajectory fixtures).
"""
from . import tensorly_np  # noqa: F401
