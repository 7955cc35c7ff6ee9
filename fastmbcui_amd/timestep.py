"""Time-stepped solves of a scene in rigid motion (matrix-free points form).

The reference's drivers (simuspclf1d_mfs_coins.py and the KTAB class,
core/KTAB_Mfs_spclf1d.py:97-128) rebuild M from the moved geometry and call
MinvF_float_fastGrad once per step, from a zero start, with the ordering and
pattern recomputed inside every call.  Here each step

1. moves the bodies (host, numpy) and uploads points, normals and sources;
2. re-plans (host maximin ordering, GPU grid pattern) every `replan_every`
   steps -- the plan is only the sparsity structure, so keeping it for a few
   steps of small motion changes the preconditioner, not the solution;
3. refactors from the points (FP32 MFMA Gram, FP64 retry of failed pivots);
4. solves matrix-free with FP64 refinement to `tol`, warm-started from the
   previous step's strengths (the fallback and judge: every step meets the
   same residual tolerance, the guess only changes the iteration count).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .scenes import RigidScene


@dataclass
class StepInfo:
    step: int
    plan_ms: float
    factor_ms: float
    solve_ms: float
    total_ms: float
    inner_iters: int
    sweeps: int
    residual: float
    replanned: bool
    warm: bool
    extra: dict = field(default_factory=dict)


class TimeStepper:
    def __init__(self, rscene: RigidScene, rho: float = 4.0, tol: float = 1e-8, device: int = 0,
                 warm: bool = True, replan_every: int = 1, precision: int = _lib.MFS_FP32,
                 maxit: int = 2000, inner_tol: float = 0.5, check_residual: bool = False):
        import torch
        self.torch = torch
        self.rs = rscene
        self.rho = float(rho)
        self.tol = float(tol)
        self.warm = bool(warm)
        self.replan_every = max(int(replan_every), 1)
        self.precision = precision
        self.maxit = int(maxit)
        self.inner_tol = float(inner_tol)
        self.check_residual = check_residual
        self.dev = torch.device("cuda", device)
        self.ctx = _lib.Context(device)
        self.X = None
        self.k = 0

    def step(self, dt: float | None = None) -> StepInfo:
        """Advance by dt (none for the first call, which solves the initial pose)."""
        torch = self.torch
        t0 = time.perf_counter()
        if dt is not None:
            self.rs.advance(dt)
        sc = self.rs.scene()
        x = torch.from_numpy(sc.x).to(self.dev)
        nx = torch.from_numpy(sc.nx).to(self.dev)
        y = torch.from_numpy(sc.y).to(self.dev)
        B = torch.from_numpy(sc.b.astype(np.float32)).to(self.dev)
        replan = (self.k % self.replan_every) == 0
        plan_ms = 0.0
        if replan:
            self.ctx.plan(sc.y, self.rho)
            plan_ms = self.ctx.stats().plan_ms
        self.ctx.factor_points(x, nx, y, self.precision)
        factor_ms = self.ctx.stats().factor_ms
        x0 = self.X if (self.warm and self.X is not None) else None
        # FP64 strengths: at tol 1e-8 an FP32 copy would not hold the fit
        X = self.ctx.solve_points64(B, maxit=self.maxit, tol=self.tol, inner_tol=self.inner_tol, x0=x0)
        s = self.ctx.stats()
        solve_ms = s.solve_ms
        torch.cuda.synchronize(self.dev)
        total_ms = (time.perf_counter() - t0) * 1e3
        res = float(s.relres[0])   # the refinement's FP64 residual of the returned X
        if self.check_residual:
            res = float(self.ctx.points_residual(B, X).max())
        info = StepInfo(self.k, plan_ms, factor_ms, solve_ms, total_ms, int(s.iters[0]),
                        int(s.outer_iters), res, replan, x0 is not None,
                        dict(order_ms=s.plan_order_ms, pattern_ms=s.plan_pattern_ms,
                             n_super=s.n_super, max_local=s.max_local,
                             retried=s.retried_supernodes))
        self.X = X
        self.k += 1
        return info
