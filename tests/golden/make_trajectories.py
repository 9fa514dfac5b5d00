"""Write ``trajectories.npz``: per-iteration CP-ALS / HOOI trajectories of
the reference TensorLy 0.9.0 (numpy backend), for the functions whose exact
output no closed form or reference test pins:

* ``parafac`` (init='random' and init='svd', with and without line search),
* ``initialize_cp(init='svd')``,
* ``partial_tucker`` / ``tucker`` (init='svd').

Runs only in the build container, where the reference tree is present; the
reference itself never travels -- only the arrays written here.

Run from the repository root:  python tests/golden/make_trajectories.py
"""
import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "trajectories.npz")
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not present; trajectories.npz is already committed")
    sys.path.insert(0, REF)
    import tensorly as tl  # the reference package, read from REF
    from tensorly.decomposition import parafac, tucker
    from tensorly.decomposition._cp import initialize_cp
    from tensorly_amd import synthetic

    tl.set_backend("numpy")
    out = {}

    # 64^3 float64 rank-8 CP model, 1 % noise: init='random' (seed 0), 50 sweeps, tol=0.
    x, _ = synthetic.cp_64_rank8()
    out["c64_tensor_checksum"] = np.array([x.sum(), (x * x).sum()])
    (w, f), errs = parafac(x, 8, n_iter_max=50, init="random", random_state=0, tol=0,
                           linesearch=False, return_errors=True)
    out["c64_random_errors"] = np.asarray(errs)
    for m in range(3):
        out[f"c64_random_factor{m}"] = np.asarray(f[m])
    out["c64_random_weights"] = np.asarray(w)

    # Same tensor, line search on, 30 sweeps.
    (w, f), errs = parafac(x, 8, n_iter_max=30, init="random", random_state=0, tol=0,
                           linesearch=True, return_errors=True)
    out["c64_ls_errors"] = np.asarray(errs)
    for m in range(3):
        out[f"c64_ls_factor{m}"] = np.asarray(f[m])

    # Small 4-way problem, init='svd', normalize_factors=True, tol stopping.
    rng = np.random.default_rng(11)
    x4 = synthetic.cp_model(np.ones(5), [rng.standard_normal((s, 5)) for s in (9, 8, 7, 6)])
    x4 = synthetic.add_noise(x4, 0.05, rng)
    out["x4"] = x4
    w0, f0 = initialize_cp(x4, 5, init="svd")
    for m in range(4):
        out[f"x4_svdinit_factor{m}"] = np.asarray(f0[m])
    (w, f), errs = parafac(x4, 5, n_iter_max=40, init="svd", tol=1e-9, normalize_factors=True,
                           return_errors=True)
    out["x4_svd_norm_errors"] = np.asarray(errs)
    out["x4_svd_norm_weights"] = np.asarray(w)
    for m in range(4):
        out[f"x4_svd_norm_factor{m}"] = np.asarray(f[m])

    # Tucker HOOI on a small Tucker model, init='svd', 10 iterations, tol=0.
    xt, _ = synthetic.tucker_256_core32(dtype=np.float64, size=24, rank=4)
    out["xt"] = xt
    (core, fac), errs = tucker(xt, rank=[4, 5, 3], n_iter_max=10, tol=0, return_errors=True)
    out["xt_errors"] = np.asarray(errs)
    out["xt_core"] = np.asarray(core)
    for m in range(3):
        out[f"xt_factor{m}"] = np.asarray(fac[m])

    np.savez_compressed(OUT, **out)
    print("wrote", OUT, sorted(out))


if __name__ == "__main__":
    main()
