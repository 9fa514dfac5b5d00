"""The C-ABI library loads and exports every symbol include/tensorly_amd.h
declares; host-side argument checks raise before any GPU work; there is no
CPU path.  No compute call is made (runs without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from conftest import ROOT


def _declared():
    with open(os.path.join(ROOT, "include", "tensorly_amd.h")) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tla_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from tensorly_amd import _lib
    lib = _lib.lib()
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.tla_abi_version() == 1


def test_binding_covers_the_header():
    """Every declared entry point has its ctypes signature in one table
    (no ad hoc bindings elsewhere), and nothing undeclared is bound."""
    from tensorly_amd import _lib
    assert _lib.exported_symbols() == _declared()


def test_workspace_queries_are_host_only():
    from tensorly_amd import _lib
    lib = _lib.lib()
    sh, _k = _lib.shape_array((512, 512, 512))
    for mode in range(3):
        for prec in (_lib.PREC_FP32, _lib.PREC_BF16, _lib.PREC_BF16X3):
            nb = lib.tla_mttkrp_workspace_size(_lib.F32, 3, sh, 32, mode, prec)
            assert 0 < nb < 64 << 20
    assert lib.tla_cp_als_workspace_size(_lib.F32, 3, sh, 32, _lib.PREC_BF16X3, 1) > 0
    assert lib.tla_cp_als_dimtree_supported(_lib.F32, 3, sh, 32, _lib.PREC_FP32) == 1
    # rank above the kernel limit is refused by the driver sizing
    assert lib.tla_cp_als_workspace_size(_lib.F32, 3, sh, 65, _lib.PREC_FP32, 0) == 0


def test_bad_arguments_return_error_codes():
    from tensorly_amd import _lib
    lib = _lib.lib()
    sh, _k = _lib.shape_array((4, 5, 6))
    pp, _keep = _lib.ptr_array([None, None, None])
    rc = lib.tla_mttkrp(_lib.F32, ctypes.c_void_p(0), 3, sh, pp, None, 3, 7, 0, None, None, 0, None)
    assert rc == _lib.ERR_ARG
    assert b"mode" in lib.tla_last_error()
    rc = lib.tla_gram64(_lib.F32, ctypes.c_void_p(0), 10, 100, None, None, 0, None)
    assert rc == _lib.ERR_ARG


def test_no_cpu_path():
    import tensorly_amd as tla
    X = torch.zeros((3, 4, 5))
    f = [torch.zeros((s, 2)) for s in X.shape]
    with pytest.raises(TypeError):
        tla.unfolding_dot_khatri_rao(X, (None, f), 0)
    with pytest.raises(TypeError):
        tla.parafac(X, 2, init="random")
    with pytest.raises(TypeError):
        tla.mode_dot(X, torch.zeros((2, 4)), 1)


def test_host_validation_matches_reference_errors():
    import tensorly_amd as tla
    with pytest.raises(ValueError):
        tla.khatri_rao([torch.zeros((3, 4)), torch.zeros((3, 4)), torch.zeros((3, 2))])
    with pytest.raises(ValueError):
        tla.mode_dot(torch.zeros((2, 4, 5)), torch.zeros((3, 4)), 0)
    with pytest.raises(ValueError):
        tla.mode_dot(torch.zeros((2, 4, 5)), torch.zeros((3,)), 1)
    with pytest.raises(ValueError):
        tla.mode_dot(torch.zeros((2, 4, 5)), torch.zeros((2, 4, 5)), 0)
    with pytest.raises(ValueError):
        tla.CPTensor((torch.ones(3), [torch.ones((4, 3)), torch.ones((5, 2))]))
    assert tla.validate_cp_rank((4, 5, 6), "same") == 8


def test_unsupported_features_raise():
    """sparsity has no GPU path and raises before any device work; a host
    tensor is refused as such."""
    import tensorly_amd as tla
    X = torch.zeros((3, 4, 5))
    with pytest.raises(NotImplementedError):
        tla.parafac(X, 2, sparsity=0.1)
    with pytest.raises(TypeError, match="GPU"):
        tla.parafac(X, 2, mask=torch.ones_like(X))


def test_planner_survives_extreme_shapes():
    """The host-side planners (workspace queries) on random shapes, ranks,
    dtypes and precisions, up to 2^40 entries and mode sizes of 2^24: no
    host crash (a 32-bit grid overflow once divided by zero).  Runs in a
    child process so that a crash fails this test instead of the session."""
    import subprocess
    import sys
    code = r'''
import sys
sys.path.insert(0, %r)
import numpy as np
from tensorly_amd import _lib
L = _lib.lib()
rng = np.random.default_rng(0)
vals = [1, 2, 3, 4, 7, 8, 15, 16, 17, 31, 32, 33, 48, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513,
        1000, 1024, 4096, 65536, 1 << 20, 1 << 24]
for trial in range(1500):
    nd = int(rng.integers(2, 6))
    shape = [int(rng.choice(vals)) for _ in range(nd)]
    if np.prod(np.array(shape, dtype=np.float64)) > 2 ** 40:
        continue
    R = int(rng.choice([1, 5, 16, 17, 32, 33, 48, 64, 65]))
    sh, _k = _lib.shape_array(shape)
    for dt in (0, 1, 2):
        for prec in (0, 1, 2, 3):
            for m in range(nd):
                L.tla_mttkrp_workspace_size(dt, nd, sh, R, m, prec)
            for kb, ke in [(0, k) for k in range(1, nd)] + [(k, nd) for k in range(1, nd)]:
                L.tla_mttkrp_partial_supported(dt, nd, sh, R, kb, ke, prec)
                L.tla_mttkrp_partial_workspace_size(dt, nd, sh, R, kb, ke, prec)
            L.tla_cp_als_workspace_size(dt, nd, sh, R, prec, 1)
print("ok")
''' % ROOT
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0 and "ok" in res.stdout, (res.returncode, res.stderr[-2000:])


def test_batched_launch_shape_checked_on_the_host():
    """tla_batched_cp_als rejects a launch shape outside what it is built
    for (waves per workgroup 2 / 4 / 8, reduction tile 16 / 32 / 64 indices)
    with TLA_ERR_ARG before anything reaches the device: the pointers here
    are never dereferenced."""
    from tensorly_amd import _lib
    lib = _lib.lib()
    dummy = ctypes.c_void_p(256)

    def call(waves, ktile):
        return lib.tla_batched_cp_als(
            2, *([dummy] * 10), 0, 10, 1e-7, 3, None, 4.0, 16.0, None, 0, 0, 0, None, None, None, None, None,
            *([dummy] * 6), None, waves, 256, ktile, None)

    for ktile in (0, 8, 48, 128):
        assert call(4, ktile) == _lib.ERR_ARG
        assert b"ktile" in lib.tla_last_error()
    for waves in (1, 3, 16):
        assert call(waves, 32) == _lib.ERR_ARG
        assert b"waves" in lib.tla_last_error()
