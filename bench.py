#!/usr/bin/env python3
"""ms/solve (factor + PCG) of the preconditioned MFS solve at N = 65,536 on one MI355X.

Workload: 64 ellipsoids x 1024 Fibonacci points (N = m = n = 65,536), 8 right-hand
sides from independent rigid motions, rho = 3, lambda = 1.5, tolerance 1e-6 on
||M^T(b - Mx)|| / ||M^T b|| (recomputed in FP64 for every column), seed 0.

A step is one solve of the whole 8-column batch with M resident in HBM:
numeric factor + PCG.  Two forms run on the same M and the same plan:

* chosen: FP32 MFMA supernode factor (FP64 retry of failed pivots) + batched
  FP32 PCG over all 8 columns + FP64 iterative refinement -> `value`;
* plain:  FP64 factor + one FP64 PCG per column in turn (the drop-in, the
  reference's column loop) -> `plain_ms_per_solve`.

The plan (ordering, pattern, supernodes) is shared by both forms and reported
as `plan_ms`.  `roofline` covers the dominant kernel (the batched matvec pass
over M), timed live with HIP events around each launch on the context stream.
`cpu_baseline` runs the FP64 CPU oracle on a bounded sample (the first 8
ellipsoids of the same scene, N = 8192, all 8 right-hand sides) and scales it
to N = 65,536 by the N^2 cost of both the Gram and the dense matvec.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "ms/solve (factor+PCG) at N=65k, res≤1e-6: chosen form vs plain FP64 on MI355X"
HBM_PEAK_GBS = 8000.0


def _parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--bodies", type=int, default=64, help="ellipsoids (1024 points each)")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--other-form-steps", type=int, default=1,
                    help="steps of the chosen form with the other normal-matvec form (two passes "
                         "when the default ran one, else one), for comparison")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--tune-factor", type=int, default=1,
                    help="autotune the factor's launch shape before the timed steps (0: defaults)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--inner-tol", type=float, default=0.5)
    ap.add_argument("--points-steps", type=int, default=1,
                    help="timed solves of the matrix-free points form (0 skips it)")
    ap.add_argument("--learned-samples-every", type=int, default=10,
                    help="learned: solve every k-th pose of the 2000-step training trajectories")
    ap.add_argument("--learned-epochs", type=int, default=50)
    ap.add_argument("--workload", choices=["ellipsoids64", "timestep", "batch", "learned", "caller"],
                    default="ellipsoids64",
                    help="timestep: 128 mixed bodies x 2048 points (N=262,144) in rigid motion, "
                         "rho 4, tol 1e-8, refactor every step, warm start (points form); "
                         "batch: 256 uneven scenes (2-32 bodies x 256-2048 points), one batch")
    ap.add_argument("--caller-bodies", type=int, default=6,
                    help="caller: bodies of 10,242 vertices in the large case")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-sequential", type=int, default=1,
                    help="batch: also solve the scenes one after another (0 skips)")
    ap.add_argument("--ts-steps", type=int, default=4, help="timestep: steps after the initial solve")
    ap.add_argument("--ts-bodies", type=int, default=128)
    ap.add_argument("--ts-points", type=int, default=2048)
    ap.add_argument("--ts-dt", type=float, default=1e-2)
    ap.add_argument("--replan-every", type=int, default=1)
    ap.add_argument("--ts-maxit", type=int, default=4000, help="timestep: inner PCG iteration cap per step")
    ap.add_argument("--copy-gb", type=float, default=0.0,
                    help="run one plain device copy of this many GB first (counter calibration)")
    return ap.parse_args()


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_info():
    model, sockets = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("physical id"):
                    sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model, "sockets": max(len(sockets), 1)}


def _median_ms(fn, reps=5, warmup=1):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(sorted(ts)[len(ts) // 2])


def cpu_baseline(tol, iters_c2):
    """The FP64 CPU oracle (C, OpenMP) on this box's host cores.

    * Config 1 (8 coins, N = 4096, 1 RHS): plan + factor + PCG timed in full,
      1 warm-up + 5 repeats, median.
    * Config 2 (the bench scene, N = 65,536, 8 RHS): factor + ms/iteration x
      measured iterations.  Its dense FP64 M (34 GB, twice with the permuted
      copy) is not built; the Gram and the matvec cost scale linearly in the
      row count m for fixed n, so the oracle's factor and PCG iteration are
      timed on 2048 and 4096 rows (Gaussian values: the cost does not depend
      on them) with all 65,536 columns and the scene's full plan, and
      extrapolated linearly to m = 65,536.  Iteration
      counts are the plain FP64 form's (the same PCG as the oracle's).  The
      plan is timed in full and reported beside it.
    """
    import numpy as np

    import oracle
    from fastmbcui_amd import scenes
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    info = _cpu_info()

    # ---- config 1, in full
    c1 = scenes.coins(seed=0, k=1)
    M1 = scenes.kernel_matrix(c1.x, c1.nx, c1.y).astype(np.float32).astype(np.float64)
    B1 = c1.b.astype(np.float32).astype(np.float64)
    res1 = {}

    def run_c1():
        res1["r"] = oracle.minvf(M1, B1, c1.y, 3.0, 1000, tol)
    c1_ms = _median_ms(run_c1)

    # ---- config 2: plan in full, factor and iterations on row samples
    sc = scenes.ellipsoids(seed=0, k=8)
    t0 = time.perf_counter()
    pl = oracle.plan(sc.y, 3.0)
    plan_ms = (time.perf_counter() - t0) * 1e3
    rows = (2048, 4096)
    fac, itr = [], []
    rng = np.random.default_rng(0)
    for mr in rows:
        # The work of both depends on the plan and the row count only, not on
        # M's values; a few thousand rows of the scene's own M leave supernode
        # Grams singular (most sources see no nearby target), so the samples
        # are Gaussian rows (SPD Grams), column-major as the oracle takes them.
        Mp = rng.standard_normal((sc.n, mr))
        t1 = time.perf_counter()
        val = oracle.factor(Mp, mr, pl)
        fac.append((time.perf_counter() - t1) * 1e3)
        b = rng.standard_normal(mr)
        t1 = time.perf_counter()
        oracle.pcg(Mp, mr, pl, val, b, 0, 3, 0.0)
        itr.append((time.perf_counter() - t1) * 1e3 / 3.5)   # M^T b (half a matvec) + 3 iterations
        del Mp
    m_full = sc.m

    def at_full(t):
        slope = (t[1] - t[0]) / (rows[1] - rows[0])
        return t[1] + slope * (m_full - rows[1])
    factor_ms = at_full(fac)
    iter_ms = at_full(itr)
    its = list(iters_c2)
    per_solve = (factor_ms + sum(i + 0.5 for i in its) * iter_ms) / len(its)
    return {
        "value": round(per_solve, 1),
        "unit": "ms/solve",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle (C, FP64, OpenMP, {threads} threads) on the bench scene (N={sc.n}, 8 RHS, "
                   f"rho 3, tol {tol:g}): factor + ms/iteration x iterations, factor and iteration "
                   f"timed on 2048 and 4096 Gaussian rows (all {sc.n} columns, the scene's plan) and "
                   f"extrapolated linearly in m to {m_full}; iterations {its} from the plain FP64 form"),
        "cpu": info,
        "config2": {"factor_ms": round(factor_ms, 1), "ms_per_iteration": round(iter_ms, 2),
                    "iterations": its, "plan_ms": round(plan_ms, 1),
                    "ms_per_solve_with_plan": round(per_solve + plan_ms / len(its), 1),
                    "sample_factor_ms": [round(v, 1) for v in fac],
                    "sample_iteration_ms": [round(v, 2) for v in itr]},
        "config1": {"workload": "8 coins x 512 points (N=4096), 1 RHS, rho 3, tol %g" % tol,
                    "ms_per_solve_with_plan": round(c1_ms, 1), "iterations": res1["r"].iters,
                    "timing": "1 warm-up + 5 repeats, median"},
    }


def timestep_main(args):
    """Rigid-motion time stepping at N = 262,144 (points form, warm start)."""
    import numpy as np
    import torch

    from fastmbcui_amd import scenes
    from fastmbcui_amd.timestep import TimeStepper

    torch.cuda.set_device(0)
    rs = scenes.mixed_bodies(seed=0, n_bodies=args.ts_bodies, pts_per_body=args.ts_points)
    n = rs.local_x.shape[0]
    _log(f"bench: timestep scene {args.ts_bodies} bodies x {args.ts_points} points (N={n})")
    ts = TimeStepper(rs, rho=4.0, tol=1e-8, replan_every=args.replan_every, maxit=args.ts_maxit)
    first = ts.step()
    _log(f"bench: step 0 (cold) {first}")
    infos = []
    for i in range(args.ts_steps):
        infos.append(ts.step(args.ts_dt))
        _log(f"bench: step {i + 1} {infos[-1]}")
    warm_ms = float(np.mean([s.total_ms for s in infos])) if infos else first.total_ms
    line = {
        "metric": "ms/step (plan+factor+solve) at N=262k, rho 4, tol 1e-8, warm-started time stepping",
        "value": round(warm_ms, 2), "unit": "ms/step", "n_gpus": 1, "steps": len(infos), "warmup": 1,
        "ms_per_step": round(warm_ms, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{args.ts_bodies} mixed bodies (coins, ellipsoids, rounded boxes) x "
                               f"{args.ts_points} points (N={n}), rigid motion v,w~N(0,0.5^2), "
                               f"dt {args.ts_dt}, rho 4, tol 1e-8, replan every {args.replan_every}, "
                               f"points form, seed 0", "n": n},
        "cold_step": first.__dict__,
        "steps_detail": [s.__dict__ for s in infos],
        "mean_inner_iters_warm": float(np.mean([s.inner_iters for s in infos])) if infos else None,
        "max_residual": max([first.residual] + [s.residual for s in infos]),
    }
    print(json.dumps(line), flush=True)
    return 0


def batch_main(args):
    """256 independent uneven scenes as one batch (points form, mixed precision)."""
    import numpy as np
    import torch

    from fastmbcui_amd import _lib, autotune, scenes

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    scs = [scenes.batch_scene(s) for s in range(args.batch_scenes)]
    ns = [sc.n for sc in scs]
    _log(f"bench: batch of {len(scs)} scenes, N from {min(ns)} to {max(ns)}, total {sum(ns)}")
    xs = [torch.from_numpy(sc.x).to(dev) for sc in scs]
    nxs = [torch.from_numpy(sc.nx).to(dev) for sc in scs]
    ys = [torch.from_numpy(sc.y).to(dev) for sc in scs]
    Bs = [torch.from_numpy(sc.b.astype(np.float32)).to(dev) for sc in scs]
    threads = min(16, os.cpu_count() or 1)
    b = _lib.Batch(len(scs))
    b.plan([sc.y for sc in scs], 3.0, threads=threads)
    st = b.stats()
    _log(f"bench: batch plan {st.plan_ms:.0f} ms on {threads} threads")
    b.factor_points(xs, nxs, ys)            # warm-up (buffers)
    b.solve_points(Bs, tol=args.tol, chunk=4)
    tuned = autotune.tune_batch_chunk(b, Bs, candidates=(4, 8), tol=args.tol)
    chunk, use_graph = tuned["best"]
    _log(f"bench: launch tuning {tuned}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.factor_points(xs, nxs, ys)
    b.solve_points(Bs, tol=args.tol, chunk=chunk, use_graph=use_graph)
    torch.cuda.synchronize()
    batch_ms = (time.perf_counter() - t0) * 1e3
    st = b.stats()
    iters = [b.scene_stats(i).iters[0] for i in range(len(scs))]
    relres = [b.scene_stats(i).relres[0] for i in range(len(scs))]
    _log(f"bench: batch factor+solve {batch_ms:.0f} ms (factor {st.factor_ms:.0f}, solve {st.solve_ms:.0f}; "
         f"{st.factor_launches} factor launches over {st.factor_items} supernodes; chunk {chunk}, "
         f"{'graph' if use_graph else 'direct launches'}, {st.graph_replays} replays, {st.sweeps} sweeps)")
    seq_ms = None
    if args.batch_sequential:
        ctx = _lib.Context(0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for i, sc in enumerate(scs):
            ctx.plan(sc.y, 3.0)
            tp = ctx.stats().plan_ms
            ctx.factor_points(xs[i], nxs[i], ys[i], _lib.MFS_FP32)
            ctx.solve_points(Bs[i], tol=args.tol)
            t2 += tp / 1e3      # exclude the plan from the sequential factor+solve time
        torch.cuda.synchronize()
        seq_ms = (time.perf_counter() - t2) * 1e3
        _log(f"bench: sequential factor+solve of the same scenes {seq_ms:.0f} ms")
    line = {
        "metric": "ms per batch (factor+PCG) of 256 uneven scenes, res<=1e-6, points form",
        "value": round(batch_ms, 1), "unit": "ms/batch", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(batch_ms, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{len(scs)} scenes (2-32 coins/ellipsoids x 256-2048 points, log-uniform, "
                               f"seeds 0-{len(scs) - 1}), rho 3, tol {args.tol:g}, 1 RHS each",
                   "n_total": int(sum(ns)), "n_min": int(min(ns)), "n_max": int(max(ns))},
        "scenes_per_s": round(len(scs) / (batch_ms / 1e3), 1),
        "plan_ms": round(b.stats().plan_ms, 1), "plan_threads": threads,
        "factor_ms": round(st.factor_ms, 1), "solve_ms": round(st.solve_ms, 1),
        "sequential_factor_solve_ms": round(seq_ms, 1) if seq_ms else None,
        "launch": {"chunk": chunk, "graph": bool(use_graph), "replays": st.graph_replays,
                   "tuning_ms": tuned["ms"]},
        "factor_launches": st.factor_launches,
        "iters_max": int(max(iters)), "iters_mean": float(np.mean(iters)),
        "residual_max": float(max(relres)),
    }
    print(json.dumps(line), flush=True)
    return 0


def learned_main(args):
    """Learned warm start: train on coin trajectories (seeds 1000-1009), then
    compare zero, previous-step and learned starts on held-out trajectories."""
    import numpy as np
    import torch

    from fastmbcui_amd import _lib, warmstart

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    net = warmstart.Network(seed=0)
    data = warmstart.generate(net, seeds=list(range(1000, 1010)), steps=2000, every=args.learned_samples_every)
    gen_s = time.perf_counter() - t0
    _log(f"bench: {data.count} training samples (N={data.n}) in {gen_s:.1f} s")
    t1 = time.perf_counter()
    mse = warmstart.train(net, data, epochs=10, batch_samples=8, lr=1e-3)
    fit = warmstart.train_fit(net, data, epochs=args.learned_epochs, batch_samples=8, lr=3e-4, normal=True)
    train_s = time.perf_counter() - t1
    _log(f"bench: trained in {train_s:.1f} s: strength loss {mse[0]:.3f} -> {mse[-1]:.3f}, "
         f"normal-residual loss {fit[0]:.4f} -> {fit[-1]:.4f}")

    def evaluate(n_bodies, box, seeds, steps=5):
        ctx = _lib.Context(0)
        rows = {"zero": [], "previous": [], "learned": []}
        for seed in seeds:
            rs = warmstart.coin_trajectory(seed, n_bodies=n_bodies, box=box)
            prev = None
            for step in range(steps):
                if step:
                    rs.advance(1e-2)
                sc = rs.scene()
                ctx.plan(sc.y, 3.0)
                ctx.factor_points(*(torch.from_numpy(a).to(dev) for a in (sc.x, sc.nx, sc.y)), _lib.MFS_FP32)
                B = torch.from_numpy(sc.b.astype(np.float32)).to(dev)
                starts = {"zero": None, "learned": None}
                if prev is not None:
                    starts["previous"] = prev
                for name in list(starts):
                    torch.cuda.synchronize()
                    ts = time.perf_counter()
                    x0 = warmstart.guess(net, rs) if name == "learned" else starts[name]
                    X = ctx.solve_points(B, tol=1e-6, x0=x0)
                    torch.cuda.synchronize()
                    ms = (time.perf_counter() - ts) * 1e3
                    rows[name].append((ctx.stats().iters[0], ms, float(ctx.stats().relres[0])))
                    if name == "zero":
                        zero_X = X
                prev = zero_X
        return {k: {"iters_mean": float(np.mean([r[0] for r in v])), "ms_mean": float(np.mean([r[1] for r in v])),
                    "residual_max": float(max(r[2] for r in v))} for k, v in rows.items() if v}

    small = evaluate(8, 10.0, [1010, 1011])
    large = evaluate(128, 25.0, [1010])
    _log(f"bench: held-out N=4096: {small}")
    _log(f"bench: held-out N=65536: {large}")
    value = large["learned"]["ms_mean"]
    line = {
        "metric": "ms/solve to res<=1e-6 from the learned start, held-out coin scenes at N=65,536",
        "value": round(value, 2), "unit": "ms/solve", "n_gpus": 1, "steps": 5, "warmup": 0,
        "ms_per_step": round(value, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 weights, f32+f64 solve", "data": "synthetic",
        "config": {"workload": "train: 8 coins x 512 points, trajectories of 2000 steps dt 1e-2, seeds "
                               f"1000-1009, every {args.learned_samples_every}th pose solved ({data.count} "
                               f"samples); eval: held-out seeds, 128 coins x 512 points (N=65,536), rho 3"},
        "train_samples": data.count, "generate_s": round(gen_s, 1), "train_s": round(train_s, 1),
        "loss_strength": [round(mse[0], 4), round(mse[-1], 4)],
        "loss_normal_residual": [round(fit[0], 5), round(fit[-1], 5)],
        "heldout_N4096": small, "heldout_N65536": large,
    }
    print(json.dumps(line), flush=True)
    return 0


def _torch_solid_angle(V, F, S):
    """The reference's torch expression of compute_solid_angle
    (utils/torch_utils.py:306-334), run by PyTorch on the same GPU: the
    comparison point for the HIP kernel."""
    import torch
    F = F.long()
    s = S.unsqueeze(1)
    a, b, c = V[F[:, 0]] - s, V[F[:, 1]] - s, V[F[:, 2]] - s
    num = torch.einsum("ijk,ijk->ij", torch.cross(b, c, dim=2), a)
    l1, l2, l3 = a.norm(dim=2), b.norm(dim=2), c.norm(dim=2)
    den = (l1 * l2 * l3 + torch.einsum("ijk,ijk->ij", a, b) * l3
           + torch.einsum("ijk,ijk->ij", a, c) * l2 + torch.einsum("ijk,ijk->ij", b, c) * l1)
    return 2 * torch.atan2(num, den).transpose(0, 1)


def _torch_single_layer(S, C):
    """The reference's compute_single_layer (utils/torch_utils.py:254-276): the dense matrix."""
    return 1.0 / (C.unsqueeze(1) - S.unsqueeze(0)).norm(dim=2)


def _gpu_ms(fn, reps, warmup=1):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


# FP32 operations per solid-angle entry in k_solid_angle, counted from its
# expression (9 differences, 3 squared norms at 5, 3 square roots, N at 5,
# 3 dot products at 5, D at 8, atan2 and the doubling at 1 each).
SA_FLOPS_PER_ENTRY = 57
VALU_FP32_PEAK_TFLOPS = 157.3


def caller_main(args):
    """Caller steps of the Kirchhoff-tensor step (KTAB:97-128) on one MI355X:
    face solid angles M (forward, backward) and the single-layer potential
    SL @ X (forward, backward) at the shipped scene size (3 bodies of 642
    vertices / 1280 faces, nv = 18) and at a large size (`--caller-bodies`
    bodies of 10,242 vertices / 20,480 faces), the reference's torch
    expressions on the same GPU beside them, and the whole step with
    L.backward() at the shipped size."""
    import numpy as np
    import torch

    from fastmbcui_amd import _lib, caller, meshes, runtime

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    ctx = _lib.Context(0)
    out = {}

    def scene(nb, subdiv):
        Vs, Fs, Ss, ys, off = [], [], [], [], 0
        for b in meshes.bodies(nb, subdiv, seed=0, spacing=3.0):
            Vs.append(b["V0"] @ b["R"].T + b["y"])
            Ss.append(b["S0"] @ b["R"].T + b["y"])
            Fs.append(b["F"])
            ys.append(b["y"])
        return Vs, Fs, Ss, ys

    def flat(Vs, Fs, Ss):
        off, F = 0, []
        for V, Fi in zip(Vs, Fs):
            F.append(Fi + off)
            off += V.shape[0]
        g = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
        return (g(np.concatenate(Vs), torch.float32), g(np.concatenate(F).astype(np.int32), torch.int32),
                g(np.concatenate(Ss), torch.float32))

    for label, nb, subdiv, reps in (("shipped", 3, 3, 20), ("large", args.caller_bodies, 5, 3)):
        Vs, Fs, Ss, ys = scene(nb, subdiv)
        V, F, S = flat(Vs, Fs, Ss)
        nf, n = F.shape[0], S.shape[0]
        k = 6 * nb
        entries = nf * n
        C = caller.compute_face_centers(V, F)
        X = torch.randn((n, k), device=dev)
        G = torch.randn((nf, k), device=dev)
        _log(f"bench caller {label}: {nb} bodies, faces {nf}, sources {n}, k {k}")
        M = ctx.solid_angle(V, F, S)
        sa_ms = _gpu_ms(lambda: ctx.solid_angle(V, F, S), reps)
        gM = torch.randn_like(M)
        sab_ms = _gpu_ms(lambda: ctx.solid_angle_backward(V, F, S, gM), max(1, reps // 2))
        del gM, M
        torch.cuda.empty_cache()
        sl_ms = _gpu_ms(lambda: ctx.single_layer_apply(S, C, X), reps)
        slb_ms = _gpu_ms(lambda: ctx.single_layer_backward(S, C, X, G), reps)
        r = dict(faces=nf, sources=n, k=k, solid_angle_ms=round(sa_ms, 4),
                 solid_angle_backward_ms=round(sab_ms, 4), single_layer_ms=round(sl_ms, 4),
                 single_layer_backward_ms=round(slb_ms, 4),
                 solid_angle_entries_per_s=entries / (sa_ms * 1e-3),
                 solid_angle_tflops=SA_FLOPS_PER_ENTRY * entries / (sa_ms * 1e-3) / 1e12,
                 solid_angle_write_gbs=entries * 4 / (sa_ms * 1e-3) / 1e9)
        _log(f"bench caller {label}: solid angle {sa_ms:.3f} ms "
             f"({r['solid_angle_tflops']:.1f} TFLOP/s, {r['solid_angle_write_gbs']:.0f} GB/s written), "
             f"backward {sab_ms:.3f} ms; single layer {sl_ms:.3f} ms, backward {slb_ms:.3f} ms")
        if label == "shipped" or nf * n * 3 * 4 * 12 < 200e9:
            Fl = F.long()
            tsa = _gpu_ms(lambda: _torch_solid_angle(V, Fl, S), max(1, reps // 2))
            Vg, Sg = V.clone().requires_grad_(), S.clone().requires_grad_()
            gM = torch.randn((nf, n), device=dev)

            def torch_bwd():
                Vg.grad = Sg.grad = None
                (_torch_solid_angle(Vg, Fl, Sg) * gM).sum().backward()

            tsab = _gpu_ms(torch_bwd, max(1, reps // 2))
            tsl = _gpu_ms(lambda: _torch_single_layer(S, C) @ X, reps)
            r.update(torch_solid_angle_ms=round(tsa, 4), torch_solid_angle_fwd_bwd_ms=round(tsab, 4),
                     torch_single_layer_ms=round(tsl, 4))
            _log(f"bench caller {label}: reference torch expressions on this GPU: solid angle "
                 f"{tsa:.3f} ms, forward+backward {tsab:.3f} ms, single layer + matmul {tsl:.3f} ms")
            del gM, Vg, Sg
            torch.cuda.empty_cache()
        out[label] = r

    # the whole Kirchhoff step (forward + L.backward()) at the shipped size
    runtime.reset()
    Vs, Fs, Ss, ys = scene(3, 3)
    g = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to(dev, torch.float32)
    V_s, S_s, y_s = [g(v) for v in Vs], [g(s) for s in Ss], [g(y) for y in ys]
    F_s = [torch.as_tensor(f).to(dev) for f in Fs]
    Qp = torch.eye(18, device=dev)
    v = torch.linspace(-1, 1, 18, device=dev)

    def step(hip):
        shift = torch.zeros((3, 3), device=dev, requires_grad=True)
        Vm = [V + shift[i] for i, V in enumerate(V_s)]
        Sm = [S + shift[i] for i, S in enumerate(S_s)]
        ym = [y + shift[i] for i, y in enumerate(y_s)]
        kw = {} if hip else dict(solid_angle=_torch_solid_angle, single_layer=_torch_single_layer)
        # the coin driver's solver settings (simuspclf1d_mfs_coins.py:21-23)
        Mf = caller.kirchhoff_tensor(Vm, F_s, Sm, ym, Qp, 1.0, 1.2, 75, 2.7e-3, **kw)[0]
        L = 0.5 * v @ Mf @ v
        L.backward()
        return shift.grad

    def wall(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    ms_hip = wall(lambda: step(True), 5)
    solve_ms = runtime.last_info.solve_ms
    ms_torch = wall(lambda: step(False), 5)
    gh, gt = step(True), step(False)
    agree = float((gh - gt).norm() / gt.norm())
    out["step"] = dict(ms_hip=round(ms_hip, 2), ms_torch_expressions=round(ms_torch, 2),
                       grad_rel_diff=agree, last_backward_solve_ms=round(solve_ms, 2))
    _log(f"bench caller step (3 bodies, forward + L.backward()): {ms_hip:.1f} ms with the HIP "
         f"kernels, {ms_torch:.1f} ms with the reference's torch expressions (gradients agree to {agree:.1e})")

    # the whole caller step with kinematics, L.backward() and the Jacobian of
    # Mf_KDetached (KTAB:88-163): HIP kernels + one batched backward, against the
    # reference's composition (torch expressions, nv^2 backward() calls)
    from fastmbcui_amd import kinematics as kin

    def ref_step(table, bodies, q, v):
        qt = torch.as_tensor(q, dtype=torch.float32, device=dev).clone().requires_grad_(True)
        vt = torch.as_tensor(v, dtype=torch.float32, device=dev)
        R_s, y_s, Qp_s = kin.poses(table, qt)
        Qp = torch.cat(Qp_s, 0)
        V_s = [V0 @ R.T + y for (V0, _, _), R, y in zip(bodies, R_s, y_s)]
        S_s = [S0 @ R.T + y for (_, _, S0), R, y in zip(bodies, R_s, y_s)]
        Mf, Ks, _ = caller.kirchhoff_tensor(V_s, [f for _, f, _ in bodies], S_s, y_s, Qp, 1.0, 1.2, 75,
                                            2.7e-3, solid_angle=_torch_solid_angle,
                                            single_layer=_torch_single_layer)
        (0.5 * vt @ Mf @ vt).backward(retain_graph=True)
        R6n = kin.r6n(R_s)
        K = (R6n.T @ caller.iq_matrix(len(bodies), torch.float32, dev) @ Ks).detach()
        MfK = Qp.T @ R6n @ K
        J = torch.zeros((table.nv, table.nv, table.nq), device=dev)
        for i in range(table.nv):
            for j in range(table.nv):
                qt.grad = None
                MfK[i, j].backward(retain_graph=True)
                J[i, j] = qt.grad
        return J

    for nb in (2, 6):
        table = kin.JointTable([kin.Joint(kind=kin.KINDS["PlanarYZV2"]) for _ in range(nb)])
        bodies = []
        for b in range(nb):
            V0, F0 = meshes.ellipsoid((1.0, 1.0, 0.3) if b % 2 == 0 else (1.0, 0.6, 0.6), 3)
            bodies.append((torch.as_tensor(V0, dtype=torch.float32, device=dev),
                           torch.as_tensor(F0, device=dev),
                           torch.as_tensor(meshes.sources(V0, F0, 0.1), dtype=torch.float32, device=dev)))
        q = np.zeros(table.nq)
        q[0::3] = 2.6 * np.arange(nb)
        q[2::3] = 0.3
        v = np.linspace(-1, 1, table.nv)
        runtime.reset()
        ms_step = wall(lambda: caller.kirchhoff_step(table, bodies, q, v, 1.0, 1.2, 75, 2.7e-3), 3)
        runtime.set_solver_mode("mixed")
        runtime.reset()
        ms_mixed = wall(lambda: caller.kirchhoff_step(table, bodies, q, v, 1.0, 1.2, 75, 2.7e-3), 3)
        runtime.set_solver_mode("plain")
        runtime.reset()
        ms_ref = wall(lambda: ref_step(table, bodies, q, v), 2)
        Jh = caller.kirchhoff_step(table, bodies, q, v, 1.0, 1.2, 75, 2.7e-3)["grad_Mf_q_KDetached"]
        Jr = ref_step(table, bodies, q, v)
        dj = float((Jh - Jr).norm() / Jr.norm())
        out[f"ktab_step_{nb}_links"] = dict(nv=table.nv, ms_hip_batched=round(ms_step, 2),
                                            ms_hip_batched_mixed_solve=round(ms_mixed, 2),
                                            ms_reference_composition=round(ms_ref, 2),
                                            jacobian_rel_diff=dj)
        _log(f"bench caller KTAB step ({nb} planar links, nv {table.nv}, kinematics + Mf + L.backward() "
             f"+ d(Mf_KDetached)/dq): {ms_step:.1f} ms (HIP kernels, one batched backward; "
             f"{ms_mixed:.1f} ms with the mixed solve) vs {ms_ref:.1f} ms (torch expressions, "
             f"{table.nv ** 2} backward() calls); Jacobians agree to {dj:.1e}")

    # the shipped problems' models (stand-in bodies of the meshes' sizes and
    # extents) at the drivers' q_init and solver settings
    from fastmbcui_amd import problems
    for name in sorted(problems.PROBLEMS):
        P = problems.load_problem(name)
        bodies = [(torch.as_tensor(V, dtype=torch.float32, device=dev), torch.as_tensor(F, device=dev),
                   torch.as_tensor(S, dtype=torch.float32, device=dev)) for V, F, S in P["bodies"]]
        v = np.linspace(-1, 1, P["table"].nv)
        runtime.reset()
        run = lambda: caller.kirchhoff_step(P["table"], bodies, P["q_init"], v, P["rho_fluid"],
                                            P["rho_mfs"], P["pcg_max_iter"], P["tol_r"])
        ms = wall(run, 3)
        n = sum(b[0].shape[0] for b in P["bodies"])
        nf = sum(b[1].shape[0] for b in P["bodies"])
        out[f"problem_{name}"] = dict(ms_step=round(ms, 2), sources=n, faces=nf, nv=P["table"].nv)
        _log(f"bench caller problem {name} (stand-in bodies, {nf} faces, {n} sources, nv "
             f"{P['table'].nv}): Kirchhoff step with L.backward() and d(Mf_KDetached)/dq {ms:.1f} ms")

    cpu = None
    if not args.skip_cpu:
        from oracle import caller_oracle as co
        Vs, Fs, Ss, ys = scene(3, 3)
        Vc = torch.as_tensor(np.concatenate(Vs))
        off = np.cumsum([0] + [v.shape[0] for v in Vs[:-1]])
        Fc = torch.as_tensor(np.concatenate([f + o for f, o in zip(Fs, off)]))
        Sc = torch.as_tensor(np.concatenate(Ss))
        torch.set_num_threads(16)
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 10.0:
            co.solid_angle(Vc, Fc, Sc)
            reps += 1
        per = (time.perf_counter() - t0) / reps
        cpu = dict(value=round(per * 1e3, 2), unit="ms/solid-angle matrix (3 bodies, shipped size)",
                   cores=16, kind="port",
                   sample=f"FP64 torch-on-CPU oracle of compute_solid_angle, {reps} calls in ~10 s")
    big = out["large"]
    line = {
        "metric": "caller steps around the solve: face solid angles M and single layer SL@X on MI355X",
        "value": round(big["solid_angle_ms"], 4), "unit": "ms per solid-angle matrix (large)",
        "n_gpus": 1, "steps": 3, "warmup": 1, "ms_per_step": round(big["solid_angle_ms"], 4),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic icosphere/ellipsoid bodies",
        "config": {"workload": f"{args.caller_bodies} bodies x 10,242 vertices / 20,480 faces; "
                               "3 bodies x 642 / 1280 (shipped size)"},
        "roofline": {"bound": "valu", "achieved": round(big["solid_angle_tflops"], 2),
                     "peak": VALU_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(big["solid_angle_tflops"] / VALU_FP32_PEAK_TFLOPS, 4),
                     "traffic": None},
        "cpu_baseline": cpu,
        "detail": out,
    }
    print(json.dumps(line), flush=True)
    return 0


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard target rows across GPUs, one all-reduce per pass)", file=sys.stderr)
        return 2
    if args.workload == "timestep":
        return timestep_main(args)
    if args.workload == "batch":
        return batch_main(args)
    if args.workload == "learned":
        return learned_main(args)
    if args.workload == "caller":
        return caller_main(args)

    import numpy as np
    import torch

    from fastmbcui_amd import _lib, scenes

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    _log(f"bench: building scene ({args.bodies} ellipsoids)")
    sc = scenes.ellipsoids(seed=0, n_bodies=args.bodies, k=8)
    m, n, k = sc.m, sc.n, sc.k
    ctx = _lib.Context(0)
    x = torch.from_numpy(sc.x).to(dev)
    nx = torch.from_numpy(sc.nx).to(dev)
    y = torch.from_numpy(sc.y).to(dev)
    M = ctx.assemble_dense(x, nx, y)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(dev)
    torch.cuda.synchronize()
    _log(f"bench: M assembled ({m} x {n}, {m * n * 4 / 1e9:.1f} GB FP32)")

    if args.copy_gb > 0:
        nel = int(args.copy_gb * 1e9 / 4)
        src = torch.ones(nel, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        _log(f"bench: calibration copy of {nel * 4} bytes done")
        del src, dst

    ctx.plan(sc.y, 3.0)
    st = ctx.stats()
    plan_ms = st.plan_ms
    _log(f"bench: plan {plan_ms:.1f} ms, supernodes {st.n_super}, max local {st.max_local}, "
         f"nnz(L) {st.nnz_factor}")

    # factor launch shape (tile-count buckets, transpose ranges, streams),
    # tuned on this problem before the timed steps; the context is re-planned
    # with the winner (the plan itself is not part of the metric)
    from fastmbcui_amd import autotune
    ftune = None
    if args.tune_factor:
        ftune = autotune.tune_factor(ctx, M, sc.y, 3.0, _lib.MFS_FP32)
        _log(f"bench: factor tuning best {ftune['best']}, splits {ftune['splits']}: " +
             ", ".join(f"{c}: {t:.2f} ms" for c, t in sorted(ftune["ms"].items(), key=lambda kv: kv[1])) +
             f"; splits {ftune['split_ms']}")

    # ---------------------------------------------------------------- chosen form
    def chosen():
        ctx.factor_dense(M, _lib.MFS_FP32)
        f_ms = ctx.stats().factor_ms
        X = ctx.solve_mixed(M, B, maxit=1000, tol=args.tol, inner_tol=args.inner_tol)
        return X, f_ms

    for i in range(args.warmup):
        chosen()
        _log(f"bench: warmup {i + 1}/{args.warmup} done")
    classes = [_lib.MFS_K_MATVEC_ROWS, _lib.MFS_K_MATVEC_COLS, _lib.MFS_K_GRAM,
               _lib.MFS_K_RESID_ROWS, _lib.MFS_K_RESID_COLS, _lib.MFS_K_MATVEC_FUSED]
    ctx.track(classes)
    torch.cuda.synchronize()
    factor_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        X, f_ms = chosen()
        factor_ms.append(f_ms)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    s = ctx.stats()
    iters = [int(s.iters[c]) for c in range(k)]
    outer = int(s.outer_iters)
    retried = int(s.retried_supernodes)
    tracked = {c: ctx.tracked(c) for c in classes}
    ctx.track([])
    res = ctx.normal_residual(M, B, X)
    _log(f"bench: chosen {ms_per_step:.2f} ms/step, factor {np.mean(factor_ms):.2f} ms, inner iters {iters}, "
         f"outer {outer}, max residual {res.max():.2e}")

    # dominant matvec class (largest total event time)
    names = {_lib.MFS_K_MATVEC_ROWS: "k_mv_rows_f32x8 (y = M p, FP32, 8 RHS)",
             _lib.MFS_K_MATVEC_COLS: "k_mv_cols_f32x8 (partial M^T y, FP32, 8 RHS)",
             _lib.MFS_K_RESID_ROWS: "k_mv_rows_f64x8 (y = b - M x, FP64, 8 RHS)",
             _lib.MFS_K_RESID_COLS: "k_mv_cols_f64x8 (partial M^T y, FP64, 8 RHS)",
             _lib.MFS_K_MATVEC_FUSED: "k_mv_fused8 (q = M^T M p in one pass over M, FP32, 8 RHS)"}
    # The roofline's unit is the PCG's normal matvec q = M^T (M p), whose
    # algorithmic traffic is one pass over M (m n 4 bytes): the one-pass kernel
    # when it runs, else the pair of passes (row pass + column pass, the same
    # number of launches each), timed live per launch.
    mv_bytes = float(m) * float(n) * 4.0
    f_l, f_ms, _ = tracked[_lib.MFS_K_MATVEC_FUSED]
    r_l, r_ms, _ = tracked[_lib.MFS_K_MATVEC_ROWS]
    c_l, c_ms, _ = tracked[_lib.MFS_K_MATVEC_COLS]
    if f_l > 0:
        dom_name = names[_lib.MFS_K_MATVEC_FUSED]
        launches, avg_ms = f_l, f_ms / f_l
        dom_key = "k_mv_fused8"
    else:
        dom_name = ("normal matvec q = M^T (M p), FP32, 8 RHS: k_mv_rows_f32x8 + k_mv_cols_f32x8 "
                    "(two passes over M)")
        launches = min(r_l, c_l)
        avg_ms = (r_ms / max(r_l, 1) + c_ms / max(c_l, 1))
        dom_key = "normal_matvec"
    achieved = mv_bytes / (avg_ms * 1e-3) / 1e9 if launches else 0.0
    per_pass = None
    if f_l == 0 and r_l and c_l:
        per_pass = {"k_mv_rows_f32x8_avg_ms": round(r_ms / r_l, 4),
                    "k_mv_cols_f32x8_avg_ms": round(c_ms / c_l, 4),
                    "pass_stream_gbs": [round(mv_bytes / (r_ms / r_l * 1e-3) / 1e9, 1),
                                        round(mv_bytes / (c_ms / c_l * 1e-3) / 1e9, 1)]}
    g_l, g_ms, g_bytes = tracked[_lib.MFS_K_GRAM]
    fused_launches = tracked[_lib.MFS_K_MATVEC_FUSED][0]
    kernel_ms = {names.get(c, "supernode factor (Grams + Cholesky + L, side streams overlapped)"): round(tracked[c][1] / args.steps, 3)
                 for c in classes}

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(dom_key)
        except Exception:
            traffic = None

    # The same step with the other form of the normal matvec, for comparison:
    # two passes when the default ran the one-pass kernel (k_fused.hip), else
    # the one-pass kernel forced.
    other_form = None
    if args.other_form_steps > 0:
        ran_fused = fused_launches > 0
        ctx.set_param(_lib.MFS_PARAM_FUSED, 1 if ran_fused else 2)
        chosen()
        torch.cuda.synchronize()
        oc = [_lib.MFS_K_MATVEC_ROWS, _lib.MFS_K_MATVEC_COLS, _lib.MFS_K_MATVEC_FUSED]
        ctx.track(oc)
        t2 = time.perf_counter()
        for i in range(args.other_form_steps):
            X2, _ = chosen()
        torch.cuda.synchronize()
        other_ms = (time.perf_counter() - t2) * 1e3 / args.other_form_steps
        ot = {c: ctx.tracked(c) for c in oc}
        ctx.track([])
        s2 = ctx.stats()
        res2 = ctx.normal_residual(M, B, X2)
        ctx.set_param(_lib.MFS_PARAM_FUSED, 0)
        if ran_fused:
            (rl, rms, _), (cl, cms, _) = ot[oc[0]], ot[oc[1]]
            mv = {"k_mv_rows_f32x8_avg_ms": round(rms / max(rl, 1), 4),
                  "k_mv_cols_f32x8_avg_ms": round(cms / max(cl, 1), 4)}
        else:
            fl, fms, _ = ot[oc[2]]
            mv = {"k_mv_fused8_avg_ms": round(fms / max(fl, 1), 4)}
        other_form = dict(form=("two passes (k_mv_rows_f32x8 + k_mv_cols_f32x8)" if ran_fused
                                else "one pass (k_mv_fused8)"),
                          ms_per_solve=round(other_ms / k, 4), iters=[int(s2.iters[c]) for c in range(k)],
                          residual_max=float(res2.max()),
                          max_rel_diff_x=float((X2.double() - X.double()).norm() / X.double().norm()), **mv)
        _log(f"bench: other matvec form ({other_form['form']}) {other_ms:.2f} ms/step, {mv}, "
             f"iters {other_form['iters']}")

    # ---------------------------------------------------------------- plain form
    plain = None
    if not args.skip_plain:
        ctx.factor_dense(M, _lib.MFS_FP64)   # warm the FP64 buffers
        torch.cuda.synchronize()
        ctx.track(classes)
        tp = []
        for i in range(args.plain_steps):
            t2 = time.perf_counter()
            ctx.factor_dense(M, _lib.MFS_FP64)
            pf = ctx.stats().factor_ms
            Xp = ctx.solve_dense64(M, B, maxit=1000, tol=args.tol)
            torch.cuda.synchronize()
            tp.append((time.perf_counter() - t2) * 1e3)
        sp = ctx.stats()
        ptr = {c: ctx.tracked(c) for c in classes}
        ctx.track([])
        pres = ctx.normal_residual(M, B, Xp.float())
        err_plain = [float((X[:, c].double() - Xp[:, c]).norm() / Xp[:, c].norm()) for c in range(k)]
        pnames = {_lib.MFS_K_MATVEC_ROWS: "k_mv_rows (y = M p, FP64, 1 RHS)",
                  _lib.MFS_K_MATVEC_COLS: "k_mv_cols (partial M^T y, FP64, 1 RHS)",
                  _lib.MFS_K_GRAM: "supernode factor FP64"}
        plain = dict(ms_per_solve=float(np.mean(tp)) / k, factor_ms=pf,
                     iters=[int(sp.iters[c]) for c in range(k)], residual_max=float(pres.max()),
                     chosen_rel_err_vs_plain_fp64=max(err_plain),
                     kernel_avg_ms={pnames[c]: round(ptr[c][1] / max(ptr[c][0], 1), 4) for c in pnames})
        _log(f"bench: plain {np.mean(tp):.1f} ms/step, factor {pf:.1f} ms, iters {plain['iters']}, "
             f"max residual {pres.max():.2e}")

    # ---------------------------------------------------------------- points form
    points = None
    if args.points_steps > 0:
        pclasses = [_lib.MFS_K_MATVEC_ROWS, _lib.MFS_K_MATVEC_COLS, _lib.MFS_K_GRAM,
                    _lib.MFS_K_RESID_ROWS, _lib.MFS_K_RESID_COLS]

        # the tuned launch shape is the dense form's: the points form's Gram
        # evaluates every padded column, so it keeps the default (all tile
        # counts) and is re-planned with it (untimed)
        for prm in (_lib.MFS_PARAM_TILE_SET, _lib.MFS_PARAM_FACTOR_PHASES,
                    _lib.MFS_PARAM_FACTOR_STREAMS, _lib.MFS_PARAM_FACTOR_SPLITS):
            ctx.set_param(prm, 0)
        ctx.plan(sc.y, 3.0)

        def points_step():
            ctx.factor_points(x, nx, y, _lib.MFS_FP32)
            f_ms = ctx.stats().factor_ms
            return ctx.solve_points(B, maxit=1000, tol=args.tol, inner_tol=args.inner_tol), f_ms

        points_step()   # warm-up: buffers
        torch.cuda.synchronize()
        ctx.track(pclasses)
        tp, fp = [], []
        for i in range(args.points_steps):
            t2 = time.perf_counter()
            Xq, f_ms = points_step()
            torch.cuda.synchronize()
            tp.append((time.perf_counter() - t2) * 1e3)
            fp.append(f_ms)
        sq = ctx.stats()
        qtr = {c: ctx.tracked(c) for c in pclasses}
        ctx.track([])
        qres = ctx.points_residual(B, Xq)
        qnames = {_lib.MFS_K_MATVEC_ROWS: "k_pts_rows32 (y = M p, FP32 kernel, 8 RHS)",
                  _lib.MFS_K_MATVEC_COLS: "k_pts_cols32 (M^T y, FP32 kernel, 8 RHS)",
                  _lib.MFS_K_GRAM: "supernode factor, Gram from points",
                  _lib.MFS_K_RESID_ROWS: "k_pts_rows64 (b - M x, FP64 kernel)",
                  _lib.MFS_K_RESID_COLS: "k_pts_cols64 (M^T y, FP64 kernel)"}
        kavg = {}
        for c in pclasses:
            l, ms, fl = qtr[c]
            if l:
                kavg[qnames[c]] = {"avg_ms": round(ms / l, 4),
                                   "tflops": round(fl / l / (ms / l * 1e-3) / 1e12, 2)}
        points = dict(ms_per_solve=round(float(np.mean(tp)) / k, 4), factor_ms=round(float(np.mean(fp)), 3),
                      iters=[int(sq.iters[c]) for c in range(k)], refinement_sweeps=int(sq.outer_iters),
                      residual_max=float(qres.max()), kernels=kavg,
                      note="FP32 VALU peak 157.3 TFLOP/s (packed); these kernels issue unpacked FP32")
        _log(f"bench: points {np.mean(tp):.1f} ms/step, factor {np.mean(fp):.1f} ms, iters {points['iters']}, "
             f"max residual {qres.max():.2e}, kernels {kavg}")

    cpu = None
    if not args.skip_cpu:
        _log("bench: CPU oracle baseline (config 1 in full; config 2 by row samples)")
        cpu = cpu_baseline(args.tol, plain["iters"] if plain else iters)

    value = ms_per_step / k
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.bodies} ellipsoids x 1024 Fibonacci points (N={n}), {k} RHS, "
                               f"rho 3, lambda 1.5, tol {args.tol:g}, seed 0; chosen form",
                   "n": n, "m": m, "k": k},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": dom_name, "avg_ms": round(avg_ms, 4), "launches": launches,
                     "algorithmic_bytes": mv_bytes, "per_pass": per_pass},
        "cpu_baseline": cpu,
        "solves_per_s": round(1e3 / value, 2),
        "plan_ms": round(plan_ms, 2),
        "factor_ms": round(float(np.mean(factor_ms)), 3),
        "factor_retried_supernodes": retried,
        "factor_tuning": ({"best": {"tile_set": hex(ftune["best"][0]), "column_ranges": ftune["best"][1],
                                    "streams": ftune["best"][2],
                                    "gram_from": "bands of M" if ftune["best"][3] == 0 else "M^T",
                                    "row_splits": ftune["splits"]},
                           "ms": {f"{hex(c[0])}/{c[1]}/{c[2]}/{'bands' if c[3] == 0 else 'Mt'}": round(t, 3)
                                  for c, t in ftune["ms"].items()},
                           "row_split_ms": {str(k): round(v, 3) for k, v in ftune["split_ms"].items()}}
                          if ftune else None),
        "pcg_inner_iters": iters,
        "refinement_sweeps": outer,
        "residual_max": float(res.max()),
        "kernel_ms_per_step": kernel_ms,
        "gram_gb_gathered_per_step": round(g_bytes / max(args.steps, 1) / 1e9, 2),
        "other_matvec_form": other_form,
        "plain": plain,
        "speedup_vs_plain": (round(plain["ms_per_solve"] / value, 2) if plain else None),
        "points": points,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
