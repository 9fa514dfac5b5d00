"""The allocation-free MTTKRP (tla_mttkrp_noalloc) against tla_mttkrp with a
caller workspace, through the C ABI: the same kernels and plan, so the
outputs must agree bit for bit.  Two calls in a row, the second on a smaller
problem (the shared scratch array keeps the first call's contents), and a
problem whose scratch need exceeds the array is refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import tensorly_np as T

pytestmark = pytest.mark.gpu


def _call_both(cuda, shape, rank, mode, prec_name, dtype=torch.float32, seed=0):
    from tensorly_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(seed)
    X = torch.tensor(rng.standard_normal(shape), dtype=dtype, device=cuda)
    facs = [torch.tensor(rng.standard_normal((s, rank)), dtype=dtype, device=cuda) for s in shape]
    w = torch.tensor(rng.random(rank) + 0.5, dtype=dtype, device=cuda)
    code = _lib.dtype_code(X)
    prec = _lib.PRECISIONS[prec_name] if dtype != torch.float64 else _lib.PREC_VALU
    sh, _ks = _lib.shape_array(shape)
    pp, _kf = _lib.ptr_array([None if m == mode else f for m, f in enumerate(facs)])
    st = _lib.stream(cuda)
    nb = L.tla_mttkrp_workspace_size(code, len(shape), sh, rank, mode, prec)
    ws = torch.empty(max(int(nb), 1), dtype=torch.uint8, device=cuda)
    ref = torch.empty((shape[mode], rank), dtype=dtype, device=cuda)
    _lib.check(L.tla_mttkrp(code, _lib.ptr(X), len(shape), sh, pp, _lib.ptr(w), rank, mode, prec, _lib.ptr(ref),
                            _lib.ptr(ws), int(nb), st), "mttkrp")
    out = torch.full((shape[mode], rank), float("nan"), dtype=dtype, device=cuda)
    _lib.check(L.tla_mttkrp_noalloc(code, _lib.ptr(X), len(shape), sh, pp, _lib.ptr(w), rank, mode, prec,
                                    _lib.ptr(out), st), "mttkrp_noalloc")
    torch.cuda.synchronize()
    oracle = T.unfolding_dot_khatri_rao(X.double().cpu().numpy(), (w.double().cpu().numpy(),
                                        [f.double().cpu().numpy() for f in facs]), mode)
    return ref, out, oracle, int(nb)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_noalloc_matches_workspace_entry_two_calls(cuda, prec):
    from tensorly_amd import _lib
    cap = _lib.lib().tla_mttkrp_noalloc_max_scratch()
    assert cap == 64 << 20
    # first call: a last-mode MTTKRP with a large partial-sum scratch
    ref, out, oracle, nb = _call_both(cuda, (96, 128, 256), 32, 2, prec, seed=1)
    assert 0 < nb <= cap
    assert torch.equal(out, ref)
    assert np.linalg.norm(out.double().cpu().numpy() - oracle) / np.linalg.norm(oracle) < 1e-5
    # second call, smaller, other mode: the scratch still holds the first
    # call's partial sums
    ref2, out2, oracle2, nb2 = _call_both(cuda, (20, 24, 64), 8, 1, prec, seed=2)
    assert nb2 < nb
    assert torch.equal(out2, ref2)
    assert np.linalg.norm(out2.double().cpu().numpy() - oracle2) / np.linalg.norm(oracle2) < 1e-5


def test_noalloc_float64_and_refusal(cuda):
    from tensorly_amd import _lib
    ref, out, oracle, _ = _call_both(cuda, (17, 19, 23), 5, 0, "fp32", dtype=torch.float64, seed=3)
    assert torch.equal(out, ref)
    assert np.linalg.norm(out.cpu().numpy() - oracle) / np.linalg.norm(oracle) < 1e-12
    # a plan needing more scratch than the array holds is refused on the host
    L = _lib.lib()
    shape = (64, 64, 1 << 20)
    sh, _ks = _lib.shape_array(shape)
    need = L.tla_mttkrp_workspace_size(_lib.F32, 3, sh, 64, 2, _lib.PREC_FP32)
    assert need > L.tla_mttkrp_noalloc_max_scratch()
    pp, _kf = _lib.ptr_array([None, None, None])
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = L.tla_mttkrp_noalloc(_lib.F32, dummy, 3, sh, pp, None, 64, 2, _lib.PREC_FP32, dummy, _lib.stream(cuda))
    assert rc == _lib.ERR_ARG
    assert b"scratch" in L.tla_last_error()
