"""Autotuner for the launch shapes of the solve.

The shapes that decide how work is spread over the chip depend on the problem
(m, n, right-hand sides, scene mix), so they are measured on the problem at
hand rather than fixed:

* MFS_PARAM_POINTS_BLOCKS -- blocks per matrix-free pass (owners x splits of
  the other side): more splits fill the CUs and balance XCDs, fewer keep the
  partial sums small;
* the batch's PCG chunk -- iterations per round: longer chunks cost fewer
  polls, shorter ones launch fewer no-op kernels after a scene converged;
* the supernode factor's launch shape (MFS_PARAM_TILE_SET, _FACTOR_PHASES,
  _FACTOR_STREAMS, _FACTOR_BANDS) -- which tile counts form buckets (a
  supernode of ns rows runs in the smallest enabled 16 T >= ns: fewer buckets
  mean fewer, fuller launches but more zero padding), whether the dense FP32
  Grams come from row-major M by bands or from M^T, how many column ranges of
  the dense transpose the Gram overlaps (the M^T form only), and on how many
  streams the buckets run.

Each candidate is timed with HIP events on the work it controls; the fastest
is set and returned.  Results are kept per shape key for the process.
"""
from __future__ import annotations

import time

from . import _lib

_cache: dict = {}


def _time_ms(fn, reps):
    import torch
    fn()  # warm-up (buffers, code objects)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        fn()
        end.record()
        end.synchronize()
        best = min(best, start.elapsed_time(end))
    return best


def tune_points_blocks(ctx: _lib.Context, k: int = 8, candidates=(1024, 2048, 4096, 8192, 16384),
                       reps: int = 3) -> dict:
    """Time M p and M^T y (FP32, k columns) from the context's packed points for
    each candidate block count; keep the fastest.  Needs mfs_factor_points."""
    import torch
    key = ("points_blocks", ctx.m_points, ctx.n, k)
    if key in _cache:
        ctx.set_param(_lib.MFS_PARAM_POINTS_BLOCKS, _cache[key]["best"])
        return _cache[key]
    dev = ctx.device
    p = torch.randn(ctx.n, k, device=dev)
    y = torch.randn(ctx.m_points, k, device=dev)
    times = {}
    for c in candidates:
        ctx.set_param(_lib.MFS_PARAM_POINTS_BLOCKS, c)
        times[c] = _time_ms(lambda: (ctx.points_apply(p), ctx.points_apply(y, trans=True)), reps)
    best = min(times, key=times.get)
    ctx.set_param(_lib.MFS_PARAM_POINTS_BLOCKS, best)
    _cache[key] = {"best": best, "ms": times}
    return _cache[key]


def tune_batch_chunk(batch: _lib.Batch, Bs, candidates=(4, 8), graphs=(0, 2), maxit=1000,
                     tol=1e-6) -> dict:
    """Solve the batch once per (chunk length, launch mode) -- whole solves,
    since the chunk trades polls against launches over the whole convergence
    history; keep the fastest.  `graphs` lists the launch modes tried (0:
    direct launches; 2: per-scene graphs of `chunk` iterations)."""
    import torch
    key = ("batch_chunk", tuple(batch.n), tuple(int(b.shape[1]) for b in Bs))
    if key in _cache:
        return _cache[key]
    times = {}
    for g in graphs:
        for c in candidates:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            batch.solve_points(Bs, maxit=maxit, tol=tol, chunk=c, use_graph=g)
            times[(c, g)] = (time.perf_counter() - t0) * 1e3
    best = min(times, key=times.get)
    _cache[key] = {"best": best, "ms": {f"chunk {c}, {'graphs' if g else 'direct'}": t
                                        for (c, g), t in times.items()}}
    return _cache[key]


# Tile sets: all counts; without 3, 6 and 12 (powers of two only); without 1
# and 3 (small supernodes padded into 2 and 4 tiles).
FACTOR_TILE_SETS = (0xff, 0xd5, 0x5f)


def factor_candidates(tile_sets=FACTOR_TILE_SETS, phases=(1, 2, 4), streams=(2, 4), bands=(0, 1)):
    """(tile set, column ranges, streams, bands) tuples; bands 0 = the band
    form (column ranges do not apply: one value), 1 = the Gram from M^T."""
    out = []
    for b in bands:
        for t in tile_sets:
            for p in (phases[:1] if b == 0 else phases):
                for s in streams:
                    out.append((t, p, s, b))
    return out


def tune_factor(ctx: _lib.Context, M, xyz, rho: float, precision=_lib.MFS_FP32, candidates=None,
                reps: int = 2, points=None, splits=(1, 2, 4, 8)) -> dict:
    """Time the numeric factor of M (dense form; or from `points` = (x, nx, y)
    device tensors) for each (tile set, column ranges, streams, bands) candidate --
    re-planning for each, since the plan's buckets depend on the first two --
    and keep the fastest (the context is left planned with it); then time the
    row splits of small launches (MFS_PARAM_FACTOR_SPLITS) on that shape.
    Returns {"best": (tile_set, phases, streams, bands), "ms": {candidate: factor
    ms}, "splits": best split count, "split_ms": {splits: factor ms}}, each the
    best of `reps` timed factorisations (the context's own HIP event timing
    of the factor)."""
    key = ("factor", tuple(M.shape) if M is not None else points[0].shape[0], ctx.n, float(rho), precision)
    cands = candidates or factor_candidates()
    times = {}
    for cand in cands:
        tset, ph, st, bd = cand
        ctx.set_param(_lib.MFS_PARAM_TILE_SET, tset)
        ctx.set_param(_lib.MFS_PARAM_FACTOR_PHASES, ph)
        ctx.set_param(_lib.MFS_PARAM_FACTOR_STREAMS, st)
        ctx.set_param(_lib.MFS_PARAM_FACTOR_BANDS, bd)
        ctx.plan(xyz, rho)
        best = float("inf")
        for r in range(reps + 1):   # the first factorisation warms buffers
            if points is None:
                ctx.factor_dense(M, precision)
            else:
                ctx.factor_points(*points, precision)
            if r:
                best = min(best, ctx.stats().factor_ms)
        times[cand] = best
    win = min(times, key=times.get)
    ctx.set_param(_lib.MFS_PARAM_TILE_SET, win[0])
    ctx.set_param(_lib.MFS_PARAM_FACTOR_PHASES, win[1])
    ctx.set_param(_lib.MFS_PARAM_FACTOR_STREAMS, win[2])
    ctx.set_param(_lib.MFS_PARAM_FACTOR_BANDS, win[3])
    ctx.plan(xyz, rho)
    # then the row splits of the small launches, on the winning shape
    split_ms = {}
    for sp in splits:
        ctx.set_param(_lib.MFS_PARAM_FACTOR_SPLITS, sp)
        best = float("inf")
        for r in range(reps + 1):
            if points is None:
                ctx.factor_dense(M, precision)
            else:
                ctx.factor_points(*points, precision)
            if r:
                best = min(best, ctx.stats().factor_ms)
        split_ms[sp] = best
    best_split = min(split_ms, key=split_ms.get)
    ctx.set_param(_lib.MFS_PARAM_FACTOR_SPLITS, best_split)
    _cache[key] = {"best": win, "ms": times, "splits": best_split, "split_ms": split_ms}
    return _cache[key]
