"""Learned ALS extrapolation.

A 16 -> 64 -> 64 -> 1 MLP (tanh hidden units, linear output; 5,313
parameters) predicts the extrapolation step ``alpha`` of
``A <- A + alpha (A - A_prev)`` from 16 features of the recent error history
(the features are computed on the device by the same kernel at training and
inference time).  The output y maps to ``alpha = clip(2**y - 1, 0, alpha_max)``.

The ALS loop (native driver ``tla_cp_als_run_extrapolated``) alternates a
plain sweep with an extrapolated one; the extrapolated step is kept only if
the sweep run from the extrapolated point lowers the error below the error
before it, otherwise the saved factors are restored.  Every decision is
taken on the GPU (no host synchronisation).

Training data come from the method itself: for states of ALS trajectories
on synthetic problems, a grid of steps is probed (extrapolate, run the next
sweep, record its error) and the best step is the regression target
``log2(1 + alpha*)``.  Training runs on the GPU with the hand-written
forward / backward / Adam kernel (``tla_mlp_train``).
"""
import os

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr

__all__ = ["ExtrapolationMLP", "LearnedExtrapolation", "run_extrapolated_als", "generate_training_data",
           "N_FEATURES", "N_PARAMS", "ALPHA_GRID"]

N_FEATURES = 16
HIDDEN = 64
N_PARAMS = HIDDEN * N_FEATURES + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1
ALPHA_GRID = (0.0, 0.25, 0.5, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0, 8.0)

_sigs = False


def _bind():
    global _sigs
    if _sigs:
        return lib()
    L = lib()
    assert L.tla_mlp_num_params() == N_PARAMS
    _sigs = True
    return L


class ExtrapolationMLP:
    """Parameters (flat, 5313 floats: W1 | b1 | W2 | b2 | w3 | b3) on the GPU."""

    def __init__(self, params=None, device="cuda", seed=0):
        self.device = torch.device(device)
        if params is None:
            params = self.init_params(seed)
        self.params = torch.as_tensor(np.asarray(params, dtype=np.float32)).to(self.device).contiguous()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.t = 0

    @staticmethod
    def init_params(seed=0):
        """Glorot-uniform weights, zero biases (numpy RandomState(seed))."""
        rng = np.random.RandomState(seed)
        parts = []
        for fan_in, fan_out in ((N_FEATURES, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, 1)):
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            parts.append(rng.uniform(-lim, lim, size=(fan_out, fan_in)).ravel())
            parts.append(np.zeros(fan_out))
        return np.concatenate(parts).astype(np.float32)

    def forward(self, X):
        L = _bind()
        X = torch.as_tensor(X, dtype=torch.float32).to(self.device).contiguous()
        out = torch.empty(X.shape[0], dtype=torch.float32, device=self.device)
        check(L.tla_mlp_forward(ptr(self.params), ptr(X), X.shape[0], ptr(out), _lib.stream(self.device)),
              "mlp_forward")
        return out

    def gradient(self, X, y):
        """MSE gradient over (X, y) as one batch (len <= 128), no update."""
        L = _bind()
        X = torch.as_tensor(X, dtype=torch.float32).to(self.device).contiguous()
        y = torch.as_tensor(y, dtype=torch.float32).to(self.device).contiguous()
        order = torch.arange(X.shape[0], dtype=torch.int32, device=self.device)
        g = torch.empty_like(self.params)
        check(L.tla_mlp_train(ptr(self.params), None, None, ptr(X), ptr(y), ptr(order), X.shape[0], X.shape[0], 1,
                              0.0, 0.9, 0.999, 1e-8, 0, None, ptr(g), _lib.stream(self.device)), "mlp_train")
        return g

    def train(self, X, y, epochs=20, batch=128, lr=1e-3, seed=0, betas=(0.9, 0.999), eps=1e-8):
        """Minibatch Adam on the GPU; returns the per-step losses."""
        L = _bind()
        X = torch.as_tensor(X, dtype=torch.float32).to(self.device).contiguous()
        y = torch.as_tensor(y, dtype=torch.float32).to(self.device).contiguous()
        n = X.shape[0]
        rng = np.random.RandomState(seed)
        steps = max(1, n // batch)
        losses = []
        for _ in range(epochs):
            order = torch.as_tensor(rng.permutation(n).astype(np.int32)).to(self.device)
            loss = torch.empty(steps, dtype=torch.float32, device=self.device)
            check(L.tla_mlp_train(ptr(self.params), ptr(self.m), ptr(self.v), ptr(X), ptr(y), ptr(order), n, batch,
                                  steps, lr, betas[0], betas[1], eps, self.t, ptr(loss), None,
                                  _lib.stream(self.device)), "mlp_train")
            self.t += steps
            losses.append(loss)
        return torch.cat(losses).cpu().numpy()

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy())

    @classmethod
    def load(cls, path, device="cuda"):
        return cls(np.load(path)["params"], device=device)


class LearnedExtrapolation:
    """Extrapolation policy for :func:`tensorly_amd.parafac`: a trained
    :class:`ExtrapolationMLP`, or ``model=None`` with a fixed step."""

    def __init__(self, model=None, alpha_fixed=1.0, alpha_max=16.0, warmup=3):
        self.model = model
        self.alpha_fixed = float(alpha_fixed)
        self.alpha_max = float(alpha_max)
        self.warmup = int(warmup)


def run_extrapolated_sweeps(eng, n_sweeps, policy, features=False):
    """Run ``n_sweeps`` sweeps (plain warm-up, then plain / extrapolated
    pairs) on an engine; returns device arrays (errors, accepted, alphas[,
    features])."""
    L = _bind()
    dev = eng.dev
    errors = torch.zeros(max(n_sweeps, 1), dtype=torch.float64, device=dev)
    accepted = torch.ones(max(n_sweeps, 1), dtype=torch.int32, device=dev)
    alphas = torch.zeros(max(n_sweeps, 1), dtype=torch.float32, device=dev)
    feats = torch.zeros((max(n_sweeps, 1), N_FEATURES), dtype=torch.float32, device=dev) if features else None
    nb = L.tla_cp_als_extrap_workspace_size(eng.handle)
    ws = torch.empty(nb, dtype=torch.uint8, device=dev)
    params = policy.model.params if policy.model is not None else None
    check(L.tla_cp_als_run_extrapolated(eng.handle, int(n_sweeps), ptr(params), policy.alpha_fixed, policy.alpha_max,
                                        policy.warmup, ptr(errors), ptr(accepted), ptr(alphas), ptr(feats), ptr(ws), nb,
                                        _lib.stream(dev)), "cp_als_run_extrapolated")
    return errors, accepted, alphas, feats


def run_extrapolated_als(eng, n_iter_max, tol, policy, return_errors, normalize, cvg_criterion, callback, verbose):
    """parafac's loop with a learned (or fixed) extrapolation policy."""
    from .cp_tensor import CPTensor
    if normalize or callback is not None or tol:
        raise NotImplementedError("extrapolated ALS runs a fixed number of sweeps (tol=0, no callback, "
                                  "normalize_factors=False)")
    errors, accepted, alphas, _ = run_extrapolated_sweeps(eng, n_iter_max, policy)
    eng.raise_if_singular()
    cp = CPTensor((eng.weights.clone(), [f.clone() for f in eng.factors]))
    if return_errors:
        return cp, [float(e) for e in errors[:n_iter_max].cpu()]
    return cp


def probe_state(eng, prev_factors, alphas=ALPHA_GRID):
    """From the engine's current state (after a plain sweep; ``prev_factors``
    before it), run extrapolate + one sweep for every alpha; returns the
    errors (device FP64) and restores the state."""
    dev = eng.dev
    snap_w = eng.weights.clone()
    snap_f = [f.clone() for f in eng.factors]
    errs = torch.zeros(len(alphas), dtype=torch.float64, device=dev)
    for i, a in enumerate(alphas):
        eng.set_factors(snap_w, [f + a * (f - p) for f, p in zip(snap_f, prev_factors)])
        eng.sweeps(1, errs, i, False)
    eng.set_factors(snap_w, snap_f)
    return errs


def generate_training_data(problems, n_sweeps=24, warmup=3, alphas=ALPHA_GRID, precision="fp32", seed=0):
    """Features and targets from probed ALS trajectories.

    ``problems``: iterable of (float32 GPU tensor, rank).  For every problem a
    trajectory follows the greedy best probed step; each probed state yields
    one sample (features of the error history, target log2(1 + best alpha)).
    Returns (X [n, 16], y [n], stats)."""
    from .decomposition._cp import CpAlsEngine
    L = _bind()
    rng = np.random.RandomState(seed)
    Xs, ys = [], []
    gains = []
    for x, rank in problems:
        dev = x.device
        init = [torch.tensor(rng.random_sample((s, rank)), dtype=torch.float32, device=dev) for s in x.shape]
        eng = CpAlsEngine(x, torch.ones(rank, dtype=torch.float32, device=dev), init, precision=precision)
        errors = torch.zeros(n_sweeps + 1, dtype=torch.float64, device=dev)
        accepted = torch.ones(n_sweeps + 1, dtype=torch.int32, device=dev)
        alph = torch.zeros(n_sweeps + 1, dtype=torch.float32, device=dev)
        feats = torch.zeros((n_sweeps + 1, N_FEATURES), dtype=torch.float32, device=dev)
        probes = []
        s = 0
        for _ in range(warmup):
            eng.sweeps(1, errors, s, False)
            s += 1
        while s + 1 < n_sweeps:
            prev = [f.clone() for f in eng.factors]
            eng.sweeps(1, errors, s, False)
            check(L.tla_extrap_features(ptr(errors), ptr(alph), ptr(accepted), s, ptr(feats), _lib.stream(dev)),
                  "extrap_features")
            pe = probe_state(eng, prev, alphas)
            best = int(torch.argmin(pe))
            probes.append((s, pe))
            a = alphas[best]
            eng.set_factors(eng.weights.clone(), [f + a * (f - p) for f, p in zip(eng.factors, prev)])
            eng.sweeps(1, errors, s + 1, False)
            alph[s + 1] = a
            s += 2
        f_host = feats.cpu().numpy()
        for s_i, pe in probes:
            pe = pe.cpu().numpy()
            b = int(np.argmin(pe))
            Xs.append(f_host[s_i])
            ys.append(np.log2(1.0 + alphas[b]))
            gains.append(pe[0] / max(pe[b], 1e-300))
        eng.close()
    X = np.asarray(Xs, dtype=np.float32)
    y = np.asarray(ys, dtype=np.float32)
    return X, y, {"samples": len(y), "mean_best_alpha": float(np.mean(2.0 ** y - 1.0)),
                  "mean_error_ratio_plain_over_best": float(np.mean(gains))}


def default_model_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "extrapolation_mlp.npz")
