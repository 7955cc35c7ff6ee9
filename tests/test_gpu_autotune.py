"""GPU: the factor's launch shapes (tile-count buckets, transpose column
ranges, streams) change only how the supernodes are grouped into launches,
never the factor: every choice gives the oracle's FP64 factor to 1e-9 and a
mixed solve at the tolerance; the autotuner times each and keeps the fastest."""
import numpy as np
import pytest
import torch

import oracle
from fastmbcui_amd import _lib, autotune, scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _col_rel_err(pl, a, b):
    worst = 0.0
    for j in range(pl.n):
        s, e = pl.sptr[j], pl.sptr[j + 1]
        worst = max(worst, np.linalg.norm(a[s:e] - b[s:e]) / np.linalg.norm(b[s:e]))
    return worst


@pytest.fixture(scope="module")
def scene():
    sc = scenes.ellipsoids(seed=1, n_bodies=6, pts_per_body=512, k=3, box=10.0)
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y).astype(np.float32)
    pl = oracle.plan(sc.y, 3.0)
    ref = oracle.factor(oracle.permuted_colmajor(M.astype(np.float64), pl.Pinv), M.shape[0], pl)
    return sc, M, pl, ref


@pytest.mark.parametrize("tset,phases,streams", [(0xff, 4, 4), (0xd5, 2, 4), (0x5f, 8, 2), (0x01, 1, 1),
                                                 (0xd5, 3, 1)])
def test_factor_launch_shape_keeps_the_factor(scene, tset, phases, streams):
    sc, M, pl, ref = scene
    ctx = _lib.Context(0)
    ctx.set_param(_lib.MFS_PARAM_TILE_SET, tset)
    ctx.set_param(_lib.MFS_PARAM_FACTOR_PHASES, phases)
    ctx.set_param(_lib.MFS_PARAM_FACTOR_STREAMS, streams)
    ctx.plan(sc.y, 3.0)
    Md = torch.from_numpy(M).to(DEV)
    ctx.factor_dense(Md, _lib.MFS_FP64)
    err = _col_rel_err(pl, ctx.factor_export(), ref)
    print(f"tile set {tset:#04x}, {phases} ranges, {streams} streams: FP64 factor err {err:.2e}, "
          f"{ctx.stats().factor_ms:.2f} ms", flush=True)
    assert err < 1e-9
    # FP32 factors, FP32 and split BF16 Gram products (3e-7 .. 3e-6 measured)
    for split in (1, 2):
        ctx.set_param(_lib.MFS_PARAM_GRAM_SPLIT, split)
        ctx.factor_dense(Md, _lib.MFS_FP32)
        err32 = _col_rel_err(pl, ctx.factor_export(), ref)
        print(f"  FP32 factor (Gram products {split}) err {err32:.2e}, {ctx.stats().factor_ms:.2f} ms",
              flush=True)
        assert err32 < 5e-5, (split, err32)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    X = ctx.solve_mixed(Md, B, maxit=500, tol=1e-6)
    assert ctx.normal_residual(Md, B, X).max() <= 1e-6
    # points form under the same shape (kernel values from hi/lo FP32
    # coordinates: ~1e-14 per entry, ~1e-7 through the blocks' conditioning)
    ctx.factor_points(torch.from_numpy(sc.x).to(DEV), torch.from_numpy(sc.nx).to(DEV),
                      torch.from_numpy(sc.y).to(DEV), _lib.MFS_FP64)
    assert _col_rel_err(pl, ctx.factor_export(), oracle.factor(
        oracle.permuted_colmajor(scenes.kernel_matrix(sc.x, sc.nx, sc.y), pl.Pinv), M.shape[0], pl)) < 2e-6


def test_tune_factor_times_every_candidate(scene):
    sc, M, pl, ref = scene
    ctx = _lib.Context(0)
    Md = torch.from_numpy(M).to(DEV)
    cands = autotune.factor_candidates(phases=(2, 4), streams=(4,))
    r = autotune.tune_factor(ctx, Md, sc.y, 3.0, _lib.MFS_FP32, candidates=cands)
    print(f"factor tuning: {r}", flush=True)
    assert set(r["ms"]) == set(cands) and r["best"] in cands
    assert all(v > 0 for v in r["ms"].values())
    ctx.factor_dense(Md, _lib.MFS_FP32)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    X = ctx.solve_mixed(Md, B, maxit=500, tol=1e-6)
    assert ctx.normal_residual(Md, B, X).max() <= 1e-6


def test_factor_shape_parameters_are_checked():
    ctx = _lib.Context(0)
    for param, bad in ((_lib.MFS_PARAM_TILE_SET, 0x100), (_lib.MFS_PARAM_FACTOR_PHASES, 9),
                       (_lib.MFS_PARAM_FACTOR_STREAMS, 5), (_lib.MFS_PARAM_GRAM_SPLIT, 3)):
        with pytest.raises(_lib.MfsError):
            ctx.set_param(param, bad)


@pytest.mark.parametrize("points", [False, True])
def test_factor_row_splits_keep_the_factor(scene, points):
    """Gram rows split over up to 8 blocks per supernode (partial Grams summed
    by the last block to arrive) against one block per supernode: the same
    factor up to summation order, FP64 and FP32."""
    sc, M, pl, ref = scene
    Md = torch.from_numpy(M).to(DEV)
    pts = [torch.from_numpy(a).to(DEV) for a in (sc.x, sc.nx, sc.y)]
    vals = {}
    for splits in (1, 8):
        ctx = _lib.Context(0)
        ctx.set_param(_lib.MFS_PARAM_FACTOR_SPLITS, splits)
        ctx.plan(sc.y, 3.0)
        for prec in (_lib.MFS_FP64, _lib.MFS_FP32):
            if points:
                ctx.factor_points(*pts, prec)
            else:
                ctx.factor_dense(Md, prec)
            vals[(splits, prec)] = ctx.factor_export()
    d64 = _col_rel_err(pl, vals[(8, _lib.MFS_FP64)], vals[(1, _lib.MFS_FP64)])
    d32 = _col_rel_err(pl, vals[(8, _lib.MFS_FP32)], vals[(1, _lib.MFS_FP32)])
    e64 = _col_rel_err(pl, vals[(8, _lib.MFS_FP64)], ref)
    print(f"points={points}: split vs unsplit FP64 {d64:.2e}, FP32 {d32:.2e}; split vs oracle {e64:.2e}",
          flush=True)
    assert d64 < 1e-10
    assert d32 < 5e-2
    assert e64 < (2e-6 if points else 1e-9)
