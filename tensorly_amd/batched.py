"""Batched CP-ALS: thousands of small dense 3-way problems in one persistent,
graph-capturable launch (native entry ``tla_batched_cp_als``).

Each problem runs the CP-ALS loop of ``tensorly/decomposition/_cp.py:394-531``
(``init='random'`` factors, weights fixed at ones, ``tol`` on the absolute
change of the relative error) under one of four step policies:

* ``"plain"``      -- plain ALS;
* ``"linesearch"`` -- the reference line search (``_cp.py:333-336,
  401-403, 448-483``): every second iteration from 6 the factors jump to
  ``last + iteration**(1/acc_pow) * (current - last)``, kept if the error of
  that point is below the last recorded error, ``acc_pow`` grows after 4
  failures;
* ``"learned"``    -- the step ``A + alpha (A - A_prev)`` with ``alpha`` from
  the extrapolation network (or a fixed ``alpha``), kept only if the sweep
  run from the extrapolated point lowers the error;
* ``"probe"``      -- training-data generation: every step of a grid is
  tried at each decision and the error after one sweep recorded.

The work queue order (longest first), the workgroup size and the number of
resident workgroups are chosen by :meth:`BatchedCP.autotune`, which times
candidate launch shapes on the batch itself.
"""
import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr

__all__ = ["BatchedCP", "POLICIES", "MAX_DIM", "MAX_RANK"]

POLICIES = {"plain": 0, "linesearch": 1, "learned": 2, "probe": 3}
MAX_DIM = 64
MAX_RANK = 16
_ALIGN = 16  # elements: every packed tensor starts on a 64-byte boundary


def _cost(shape, rank):
    """Per-sweep work estimate of one problem (MFMA steps of the three
    MTTKRPs plus a solve term) for longest-first queue ordering."""
    n = int(np.prod(shape))
    c = 0
    for s in shape:
        c += ((s + 15) // 16) * (n // s)
    return c + 64 * rank


class BatchedCP:
    """A packed batch of 3-way float32 problems resident on one GPU.

    ``tensors``: sequence of 3-way arrays (numpy or torch); ``ranks``: one
    CP rank per problem."""

    def __init__(self, tensors, ranks, device=None):
        if len(tensors) != len(ranks):
            raise ValueError("one rank per tensor")
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise TypeError("BatchedCP needs a GPU device; tensorly_amd has no CPU path")
        P = len(tensors)
        shapes = np.zeros((P, 3), dtype=np.int32)
        rk = np.asarray([int(r) for r in ranks], dtype=np.int32)
        x_off = np.zeros(P, dtype=np.int64)
        f_off = np.zeros(P, dtype=np.int64)
        xo = fo = 0
        for p, t in enumerate(tensors):
            sh = tuple(int(s) for s in t.shape)
            if len(sh) != 3:
                raise ValueError(f"problem {p}: batched CP-ALS takes 3-way tensors, got shape {sh}")
            if min(sh) < 1 or max(sh) > MAX_DIM:
                raise ValueError(f"problem {p}: mode sizes must be in [1, {MAX_DIM}], got {sh}")
            if not 1 <= rk[p] <= MAX_RANK:
                raise ValueError(f"problem {p}: rank must be in [1, {MAX_RANK}], got {rk[p]}")
            shapes[p] = sh
            x_off[p] = xo
            f_off[p] = fo
            xo += -(-int(np.prod(sh)) // _ALIGN) * _ALIGN
            fo += sum(sh) * int(rk[p])
        self.n_problems = P
        self.shapes = shapes
        self.ranks = rk
        self.x_off_host = x_off
        self.f_off_host = f_off
        self.n_factor = fo
        X = torch.zeros(max(xo, _ALIGN), dtype=torch.float32, device=self.device)
        host = np.zeros(max(xo, _ALIGN), dtype=np.float32)
        for p, t in enumerate(tensors):
            a = t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)
            n = a.size
            host[x_off[p]:x_off[p] + n] = a.astype(np.float32, copy=False).reshape(-1)
        X.copy_(torch.from_numpy(host))
        self.X = X
        d = self.device
        self.x_off = torch.tensor(x_off, device=d)
        self.f_off = torch.tensor(f_off, device=d)
        self.shape_d = torch.tensor(shapes.reshape(-1), device=d)
        self.rank_d = torch.tensor(rk, device=d)
        cost = np.array([_cost(shapes[p], rk[p]) for p in range(P)], dtype=np.int64)
        self.order_host = np.argsort(-cost, kind="stable").astype(np.int32)
        self.order = torch.tensor(self.order_host, device=d)
        self.queue = torch.zeros(1, dtype=torch.int32, device=d)
        nf = max(fo, 1)
        self.f_prev = torch.zeros(nf, dtype=torch.float32, device=d)
        self.f_save = torch.zeros(nf, dtype=torch.float32, device=d)
        self.config = None

    # ------------------------------------------------------------------
    def random_init(self, seed=0):
        """Packed initial factors: problem p gets what ``init='random'`` with
        ``random_state=seed + p`` draws (``_cp.py:61-70``,
        ``random/base.py:103-105``): ``RandomState(seed + p).random_sample``
        of (I_m, R) per mode, in mode order, cast to float32."""
        host = np.zeros(max(self.n_factor, 1), dtype=np.float32)
        for p in range(self.n_problems):
            rs = np.random.RandomState(seed + p)
            off = int(self.f_off_host[p])
            r = int(self.ranks[p])
            for s in self.shapes[p]:
                f = rs.random_sample((int(s), r))
                host[off:off + f.size] = f.reshape(-1)
                off += f.size
        return torch.from_numpy(host).to(self.device)

    def factors_of(self, packed, p):
        """The three factor matrices of problem p (views into ``packed``)."""
        off = int(self.f_off_host[p])
        r = int(self.ranks[p])
        out = []
        for s in self.shapes[p]:
            out.append(packed[off:off + int(s) * r].view(int(s), r))
            off += int(s) * r
        return out

    # ------------------------------------------------------------------
    def _grid(self, waves, blocks_per_cu=None):
        per_cu = lib().tla_batched_cp_max_blocks_per_cu(int(waves))
        if per_cu < 0:
            check(_lib.ERR_HIP, "batched_cp occupancy")
        per_cu = max(1, per_cu)
        if blocks_per_cu is not None:
            per_cu = max(1, min(per_cu, int(blocks_per_cu)))
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        return max(1, min(self.n_problems, cus * per_cu))

    def make_outputs(self, max_sweeps, history=False, max_decisions=0, n_probe=0):
        d, P = self.device, self.n_problems
        out = {
            "sweeps": torch.zeros(P, dtype=torch.int32, device=d),
            "passes": torch.zeros(P, dtype=torch.int32, device=d),
            "error": torch.zeros(P, dtype=torch.float64, device=d),
            "n_extrap": torch.zeros(P, dtype=torch.int32, device=d),
            "n_accept": torch.zeros(P, dtype=torch.int32, device=d),
            "status": torch.zeros(P, dtype=torch.int32, device=d),
            "ticks": torch.zeros(P, dtype=torch.int64, device=d),
        }
        if history:
            out["hist_err"] = torch.zeros((P, max_sweeps), dtype=torch.float64, device=d)
            out["hist_alpha"] = torch.zeros((P, max_sweeps), dtype=torch.float32, device=d)
            out["hist_acc"] = torch.ones((P, max_sweeps), dtype=torch.int32, device=d)
        if max_decisions:
            out["features"] = torch.zeros((P, max_decisions, 16), dtype=torch.float32, device=d)
            out["probe_err"] = torch.zeros((P, max_decisions, n_probe), dtype=torch.float64, device=d)
        return out

    def launch(self, factors, out, policy="plain", max_sweeps=1000, tol=1e-7, warmup=3, model=None,
               alpha_fixed=0.0, alpha_max=16.0, probe_alphas=None, max_decisions=0, first_decision=0, waves=None,
               grid=None, ktile=None):
        """Enqueue one batched run on the current stream (no host sync).
        ``factors`` (packed, float32) is updated in place; ``out`` comes from
        :meth:`make_outputs`.  Safe to capture in a graph."""
        pol = POLICIES[policy]
        if pol in (2, 3) and "hist_err" not in out:
            raise ValueError(f"policy {policy!r} needs make_outputs(..., history=True)")
        if pol == 3 and (probe_alphas is None or "probe_err" not in out):
            raise ValueError("policy 'probe' needs probe_alphas and make_outputs(..., max_decisions=, n_probe=)")
        if factors.dtype != torch.float32 or factors.device != self.device or factors.numel() < self.n_factor:
            raise ValueError("factors must be the packed float32 factor vector of this batch")
        cfg = self.config or {"waves": 4, "blocks_per_cu": None, "ktile": 32}
        if waves is None or grid is None:
            waves = waves or cfg["waves"]
            grid = grid or self._grid(waves, cfg.get("blocks_per_cu"))
        ktile = int(ktile or cfg.get("ktile", 32))
        params = model.params if model is not None else None
        n_probe = 0 if probe_alphas is None else int(probe_alphas.numel())
        check(lib().tla_batched_cp_als(
            self.n_problems, ptr(self.X), ptr(self.x_off), ptr(self.shape_d), ptr(self.rank_d), ptr(factors),
            ptr(self.f_off), ptr(self.f_prev), ptr(self.f_save), ptr(self.order), ptr(self.queue), pol,
            int(max_sweeps), float(tol), int(warmup), ptr(params), float(alpha_fixed), float(alpha_max),
            ptr(probe_alphas), n_probe, int(max_decisions), int(first_decision), ptr(out.get("hist_err")), ptr(out.get("hist_alpha")),
            ptr(out.get("hist_acc")), ptr(out.get("features")), ptr(out.get("probe_err")), ptr(out["sweeps"]),
            ptr(out["passes"]), ptr(out["error"]), ptr(out["n_extrap"]), ptr(out["n_accept"]), ptr(out["status"]),
            ptr(out["ticks"]), int(waves), int(grid), ktile, _lib.stream(self.device)), "batched_cp_als")

    def run(self, factors=None, seed=0, policy="plain", max_sweeps=1000, tol=1e-7, history=False, **kw):
        """Run to convergence; returns (packed factors, outputs dict)."""
        f = self.random_init(seed) if factors is None else factors.clone()
        md = int(kw.get("max_decisions", 0))
        npb = 0 if kw.get("probe_alphas") is None else int(kw["probe_alphas"].numel())
        out = self.make_outputs(max_sweeps, history=history or policy in ("learned", "probe"),
                                max_decisions=md, n_probe=npb)
        self.launch(f, out, policy=policy, max_sweeps=max_sweeps, tol=tol, **kw)
        return f, out

    # ------------------------------------------------------------------
    def capture(self, factors, out, **kw):
        """Capture one launch (queue reset + persistent kernel) in a
        ``torch.cuda.CUDAGraph``; replay it after refilling ``factors``."""
        if kw.get("waves") is None or kw.get("grid") is None:
            cfg = self.config or {"waves": 4, "blocks_per_cu": None}
            kw["waves"] = kw.get("waves") or cfg["waves"]
            kw["grid"] = kw.get("grid") or self._grid(kw["waves"], cfg.get("blocks_per_cu"))
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            self.launch(factors, out, **kw)  # warm-up outside the graph
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch(factors, out, **kw)
        return g

    def autotune(self, candidates=None, sweeps=6, seed=0, repeats=2, converge=False, policy="plain",
                 max_sweeps=1000, tol=1e-7, **policy_kw):
        """Time candidate launch shapes (waves per workgroup x resident
        workgroups per CU x the MTTKRP's reduction tile of 16, 32 or 64
        indices per wave) and keep the fastest.  Returns the timing table.

        ``converge=False`` times ``sweeps`` plain sweeps of every problem: a
        throughput measure.  ``converge=True`` times whole runs of ``policy``
        to ``tol`` (at most ``max_sweeps``): the launch then lasts as long as
        its slowest problems, whose sweep counts vary 3-5x across a batch, and
        wider workgroups finish those sooner at some cost in throughput.
        Start from the factors of ``random_init(seed)``, so that a caller
        can tune on other starting points than the ones it times."""
        if candidates is None:
            candidates = []
            for w in (2, 4, 8):
                per_cu = lib().tla_batched_cp_max_blocks_per_cu(w)
                for b in sorted({per_cu, max(1, per_cu // 2)}):
                    for kt in (16, 32, 64):
                        candidates.append((w, b, kt))
        f0 = self.random_init(seed)
        if converge:
            pol, n_sw, tl, kw = policy, int(max_sweeps), float(tol), dict(policy_kw)
        else:
            pol, n_sw, tl, kw = "plain", int(sweeps), 0.0, {}
        out = self.make_outputs(n_sw, history=pol in ("learned", "probe"))
        table = []
        for cand in candidates:
            w, b = cand[0], cand[1]
            kt = cand[2] if len(cand) > 2 else 32
            grid = self._grid(w, b)
            f = f0.clone()
            self.launch(f, out, policy=pol, max_sweeps=n_sw, tol=tl, waves=w, grid=grid, ktile=kt, **kw)
            torch.cuda.synchronize(self.device)
            best = float("inf")
            for _ in range(repeats):
                f.copy_(f0)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                self.launch(f, out, policy=pol, max_sweeps=n_sw, tol=tl, waves=w, grid=grid, ktile=kt, **kw)
                e1.record()
                torch.cuda.synchronize(self.device)
                best = min(best, e0.elapsed_time(e1))
            table.append({"waves": w, "blocks_per_cu": b, "ktile": kt, "grid": grid, "ms": best})
        win = min(table, key=lambda t: t["ms"])
        self.config = {"waves": win["waves"], "blocks_per_cu": win["blocks_per_cu"], "ktile": win["ktile"]}
        return table
