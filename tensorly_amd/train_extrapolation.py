"""Train the ALS extrapolation network on trajectories the method generates.

Data: the batched CP-ALS kernel in its probing policy (``tla_batched_cp_als``
policy 3) runs ALS on synthetic problems; at each probed decision it records
the 16 history features and, for every step ``alpha`` of a grid, the
relative error after one sweep from ``A + alpha (A - A_prev)``.  The
regression target is ``log2(1 + alpha*)`` of the best probed step.  Three
problem families are mixed:

* the batched workload's distribution -- 3-way tensors with mode sizes
  U{8..64}, ranks U{2..16}, factor columns with pairwise congruence 0.9
  ("swamp" regime), 1 % noise;
* dense CP models with U[0,1) factors (mode sizes U{32..64}, ranks
  U{8..16}, 1 % noise), the collinearity pattern of the 512^3 rank-32
  workload at a size the batched kernel runs;
* larger U[0,1)-factor models (mode sizes U{96..256}, ranks U{16..32}),
  probed with the single-problem driver along greedy trajectories, repeated
  so that they make up about a fifth of the samples.

Probing starts at a different decision per chunk, so the samples cover the
start, the middle and the tail of the trajectories; decisions before the
probed window follow the fixed step ``alpha_fixed``.

Training: minibatch Adam with the project's forward / backward kernel
(``tla_mlp_train``) on the GPU.  The result is written to
``tensorly_amd/models/extrapolation_mlp.npz`` (``params`` plus a JSON
``meta`` record of the data and the losses).

    python -m tensorly_amd.train_extrapolation [--problems 100000]
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from . import synthetic
from .batched import BatchedCP
from .learned import ExtrapolationMLP, default_model_path

PROBE_GRID = (0.0, 0.5, 1.0, 2.0, 3.0, 4.0, 6.0, 8.0, 12.0, 16.0)


def uniform01_problems(n, seed, size_range=(32, 64), rank_range=(8, 16), noise=0.01):
    """CP models with U[0,1) factors (rng(seed): sizes, rank, factors, noise
    per problem in order)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        shape = tuple(int(s) for s in rng.integers(size_range[0], size_range[1] + 1, size=3))
        r = int(rng.integers(rank_range[0], rank_range[1] + 1))
        fs = [rng.random((s, r)) for s in shape]
        x = synthetic.add_noise(synthetic.cp_model(np.ones(r), fs), noise, rng).astype(np.float32)
        out.append((x, r))
    return out


def large_uniform01_data(n, seed, device, size_range=(96, 256), rank_range=(16, 32), n_sweeps=64, log=print):
    """Greedy probed trajectories (tensorly_amd.learned.generate_training_data)
    of U[0,1)-factor CP models built on the GPU; rng(seed) draws sizes and
    ranks, torch generator seed + i the factors and noise of problem i."""
    from .learned import generate_training_data
    rng = np.random.default_rng(seed)
    dev = torch.device(device) if device is not None else torch.device("cuda")
    Xs, ys = [], []
    t0 = time.time()
    for i in range(n):
        shape = tuple(int(v) for v in rng.integers(size_range[0], size_range[1] + 1, size=3))
        r = int(rng.integers(rank_range[0], rank_range[1] + 1))
        g = torch.Generator(device=dev)
        g.manual_seed(seed + i)
        fs = [torch.rand((s_, r), generator=g, device=dev, dtype=torch.float64) for s_ in shape]
        x = torch.einsum("ir,jr,kr->ijk", *fs)
        e = torch.randn(x.shape, generator=g, device=dev, dtype=torch.float64)
        x = (x + 0.01 * torch.linalg.vector_norm(x) / torch.linalg.vector_norm(e) * e).float().contiguous()
        X, y, _ = generate_training_data([(x, r)], n_sweeps=n_sweeps, warmup=3, alphas=PROBE_GRID, seed=seed + i)
        Xs.append(X)
        ys.append(y)
        del x, e, fs
        if (i + 1) % 25 == 0:
            log(f"[train] large problems {i + 1}/{n} in {time.time() - t0:.1f}s")
    return np.concatenate(Xs), np.concatenate(ys)


def collect(problems, first_decision, max_decisions=6, alpha_fixed=2.0, seed=0, device=None, model=None):
    """Probe one chunk of problems; returns (features [n, 16], probe errors
    [n, G], error before the step [n])."""
    b = BatchedCP([x for x, _ in problems], [r for _, r in problems], device=device)
    grid = torch.tensor(PROBE_GRID, dtype=torch.float32, device=b.device)
    max_sweeps = 3 + 2 * (first_decision + max_decisions) + 1
    f, out = b.run(seed=seed, policy="probe", max_sweeps=max_sweeps, tol=0.0, probe_alphas=grid,
                   max_decisions=max_decisions, first_decision=first_decision, alpha_fixed=alpha_fixed,
                   alpha_max=float(PROBE_GRID[-1]), model=model, warmup=3)
    feats = out["features"].cpu().numpy().reshape(-1, 16)
    pe = out["probe_err"].cpu().numpy().reshape(-1, len(PROBE_GRID))
    he = out["hist_err"].cpu().numpy()
    before = np.stack([he[:, 3 + 2 * (first_decision + d)] for d in range(max_decisions)], axis=1).reshape(-1)
    ok = (out["status"].cpu().numpy() == 0).repeat(max_decisions)
    ok &= np.all(np.isfinite(feats), axis=1) & np.all(pe < 1e299, axis=1) & (before > 0)
    return feats[ok], pe[ok], before[ok]


def targets(pe):
    """log2(1 + alpha*) of the best probed step (lowest index on ties)."""
    grid = np.asarray(PROBE_GRID)
    return np.log2(1.0 + grid[np.argmin(pe, axis=1)]).astype(np.float32)


def build_dataset(n_problems, seed=1000, chunk=4096, uniform_fraction=0.25, windows=(0, 4, 12, 30),
                  max_decisions=6, device=None, log=print):
    Xs, Ys, meta = [], [], []
    n_chunks = max(1, -(-n_problems // chunk))
    for c in range(n_chunks):
        n = min(chunk, n_problems - c * chunk)
        n_u = int(round(n * uniform_fraction))
        t0 = time.time()
        probs = synthetic.batched_cp(n_problems=n - n_u, seed=seed + 2 * c)
        probs += uniform01_problems(n_u, seed + 2 * c + 1)
        first = windows[c % len(windows)]
        f, pe, before = collect(probs, first, max_decisions=max_decisions, seed=seed + c, device=device)
        Xs.append(f)
        Ys.append(targets(pe))
        gain = float(np.mean(np.log10(np.maximum(pe[:, 0], 1e-300) / np.maximum(pe.min(axis=1), 1e-300))))
        meta.append({"chunk": c, "problems": n, "uniform01": n_u, "first_decision": first, "samples": len(f),
                     "mean_log10_gain_best_vs_plain": gain})
        log(f"[train] chunk {c + 1}/{n_chunks}: {len(f)} samples (window from decision {first}) "
            f"in {time.time() - t0:.1f}s")
    return np.concatenate(Xs), np.concatenate(Ys), meta


def train(X, y, epochs=30, batch=128, lr=2e-3, seed=0, val_fraction=0.1, device=None, log=print):
    rng = np.random.RandomState(seed)
    idx = rng.permutation(len(y))
    n_val = int(len(y) * val_fraction)
    val, tr = idx[:n_val], idx[n_val:]
    net = ExtrapolationMLP(seed=seed, device=device or "cuda")
    hist = []
    for ep in range(epochs):
        lr_ep = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))
        losses = net.train(X[tr], y[tr], epochs=1, batch=batch, lr=float(lr_ep), seed=seed + ep)
        pv = net.forward(X[val]).cpu().numpy()
        vl = float(np.mean((pv - y[val]) ** 2))
        hist.append({"epoch": ep, "train_loss": float(np.mean(losses)), "val_loss": vl})
        log(f"[train] epoch {ep}: train {hist[-1]['train_loss']:.4f} val {vl:.4f}")
    base = float(np.mean((y[val] - y[tr].mean()) ** 2))
    return net, hist, base


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--problems", type=int, default=100000)
    p.add_argument("--epochs", type=int, default=30)
    p.add_argument("--seed", type=int, default=1000)
    p.add_argument("--out", default=default_model_path())
    p.add_argument("--large", type=int, default=300, help="larger U[0,1) problems probed with the single driver")
    args = p.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    t0 = time.time()
    X, y, meta = build_dataset(args.problems, seed=args.seed, log=log)
    n_small = len(y)
    n_large = 0
    if args.large:
        XL, yL = large_uniform01_data(args.large, args.seed + 777, None, log=log)
        n_large = len(yL)
        rep = max(1, int(round(0.25 * n_small / max(1, len(yL)))))
        X = np.concatenate([X] + [XL] * rep)
        y = np.concatenate([y] + [yL] * rep)
        meta.append({"large_uniform01_problems": args.large, "samples": n_large, "repeated": rep,
                     "mean_alpha_star": float(np.mean(2.0 ** yL - 1))})
    log(f"[train] {len(y)} samples ({n_small} batched, {n_large} large) in {time.time() - t0:.1f}s; "
        f"mean alpha* {np.mean(2.0 ** y - 1):.3f}")
    net, hist, base = train(X, y, epochs=args.epochs, log=log)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    info = {"problems": args.problems, "samples": int(len(y)), "probe_grid": list(PROBE_GRID), "chunks": meta,
            "epochs": hist, "val_loss_constant_predictor": base, "seconds": time.time() - t0}
    np.savez(args.out, params=net.params.cpu().numpy(), meta=np.array(json.dumps(info)))
    log(f"[train] wrote {args.out}: val loss {hist[-1]['val_loss']:.4f} (constant predictor {base:.4f})")


if __name__ == "__main__":
    main()
