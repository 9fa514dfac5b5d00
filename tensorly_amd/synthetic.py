"""Deterministic synthetic inputs for the five benchmark workloads.

Every recipe draws from ``numpy.random.default_rng(seed)`` in a fixed order
(stated per function), so the host-side arrays are identical on any machine
up to the last-bit rounding of the BLAS product that forms the CP / Tucker
model.  Noise is scaled so that ``||noise||_F = level * ||model||_F``.

Workloads (all C-order, last mode contiguous):

* ``cp_64_rank8``     -- 64^3 float64, rank-8 CP, factors N(0,1), 1 % noise.
* ``cp_512_rank32``   -- 512^3 float32, rank-32 CP, factors U[0,1), 1 % noise.
* ``cp_128x4_rank64`` -- 128^4 float32, rank-64 CP, unit-norm N(0,1) columns,
  log-uniform weights in [1e-2, 1], 0.1 % noise.
* ``tucker_256_core32`` -- 256^3 float32, 32^3 N(0,1) core, orthonormal
  factors (Q of QR of 256x32 Gaussians), 1 % noise.
* ``batched_cp``      -- 4096 small 3-way problems, sizes U{8..64}, ranks
  U{2..16}, factor columns with pairwise congruence 0.9, 1 % noise.
"""
import numpy as np

__all__ = ["cp_model", "add_noise", "cp_64_rank8", "cp_512_rank32", "cp_128x4_rank64",
           "tucker_256_core32", "batched_cp", "collinear_factor"]


def cp_model(weights, factors, dtype=np.float64):
    """Dense ``sum_r w_r a_r o b_r o ...`` formed as one GEMM of the first
    factor against the Khatri-Rao product of the others."""
    f0 = np.asarray(factors[0], dtype=np.float64) * np.asarray(weights, dtype=np.float64)[None, :]
    kr = np.asarray(factors[1], dtype=np.float64)
    for f in factors[2:]:
        f = np.asarray(f, dtype=np.float64)
        kr = (kr[:, None, :] * f[None, :, :]).reshape(-1, kr.shape[1])
    full = f0 @ kr.T
    shape = tuple(int(f.shape[0]) for f in factors)
    return full.reshape(shape).astype(dtype, copy=False)


def add_noise(x, level, rng):
    """Return ``x + level * ||x|| * E / ||E||`` with ``E`` standard normal
    drawn in the dtype of ``x`` (float32 draws for float32 tensors)."""
    dt = np.float32 if x.dtype == np.float32 else np.float64
    e = rng.standard_normal(x.shape, dtype=dt)
    nx = float(np.linalg.norm(x.reshape(-1).astype(np.float64, copy=False)))
    ne = float(np.linalg.norm(e.reshape(-1).astype(np.float64, copy=False)))
    e *= dt(level * nx / ne)
    x = x + e
    return x


def cp_64_rank8():
    """64^3 float64; rng(0): factors A, B, C ~ N(0,1) (64x8 each, in that
    order), weights 1, then noise at 1 %.  Returns (tensor, (w, factors))."""
    rng = np.random.default_rng(0)
    factors = [rng.standard_normal((64, 8)) for _ in range(3)]
    w = np.ones(8)
    x = add_noise(cp_model(w, factors), 0.01, rng)
    return x, (w, factors)


def cp_512_rank32(dtype=np.float32):
    """512^3; rng(1): factors ~ U[0,1) (512x32 each, mode order), weights 1,
    model formed in float64 and cast, then float32 noise at 1 %."""
    rng = np.random.default_rng(1)
    factors = [rng.random((512, 32)) for _ in range(3)]
    w = np.ones(32)
    x = cp_model(w, factors, dtype=dtype)
    x = add_noise(x, 0.01, rng)
    return x, (w, factors)


def cp_128x4_rank64(dtype=np.float32, size=128, rank=64):
    """size^4 (128 by default); rng(2): four size x rank N(0,1) factors with
    unit-norm columns, weights 10**U(-2, 0), model in float64 then cast,
    0.1 % noise."""
    rng = np.random.default_rng(2)
    factors = []
    for _ in range(4):
        f = rng.standard_normal((size, rank))
        factors.append(f / np.linalg.norm(f, axis=0, keepdims=True))
    w = 10.0 ** rng.uniform(-2.0, 0.0, size=rank)
    x = cp_model(w, factors, dtype=dtype)
    x = add_noise(x, 0.001, rng)
    return x, (w, factors)


def tucker_256_core32(dtype=np.float32, size=256, rank=32):
    """size^3; rng(3): core ~ N(0,1) (rank^3), then three factors =
    Q of QR of size x rank Gaussians, model in float64, 1 % noise."""
    rng = np.random.default_rng(3)
    core = rng.standard_normal((rank, rank, rank))
    factors = [np.linalg.qr(rng.standard_normal((size, rank)))[0] for _ in range(3)]
    x = np.einsum("abc,ia->ibc", core, factors[0], optimize=True)
    x = np.einsum("ibc,jb->ijc", x, factors[1], optimize=True)
    x = np.einsum("ijc,kc->ijk", x, factors[2], optimize=True)
    x = add_noise(x.astype(dtype), 0.01, rng)
    return x, (core, factors)


def collinear_factor(rng, size, rank, congruence):
    """``size x rank`` matrix with unit columns whose pairwise cosine is
    ``congruence`` in expectation: Gaussian columns mixed by the Cholesky
    factor of the equicorrelation matrix, then normalised."""
    c = np.full((rank, rank), congruence)
    np.fill_diagonal(c, 1.0)
    L = np.linalg.cholesky(c)
    g = rng.standard_normal((size, rank))
    g /= np.linalg.norm(g, axis=0, keepdims=True)
    f = g @ L.T
    return f / np.linalg.norm(f, axis=0, keepdims=True)


def batched_cp(n_problems=4096, seed=4, size_range=(8, 64), rank_range=(2, 16),
               congruence=0.9, noise=0.01, dtype=np.float32):
    """List of ``(tensor, rank)`` 3-way problems.  rng(seed), per problem in
    order: three sizes U{lo..hi}, rank U{lo..hi}, three collinear factors,
    noise."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_problems):
        shape = tuple(int(s) for s in rng.integers(size_range[0], size_range[1] + 1, size=3))
        r = int(rng.integers(rank_range[0], rank_range[1] + 1))
        fs = [collinear_factor(rng, s, r, congruence) for s in shape]
        x = cp_model(np.ones(r), fs)
        x = add_noise(x, noise, rng).astype(dtype)
        out.append((x, r))
    return out
