"""GPU parity of the matrix-free (points) form, through the C-ABI.

* mfs_points_apply: M v and M^T v evaluated from the packed points against the
  FP64 kernel matrix (FP64 evaluation to 1e-12, FP32 hi/lo evaluation to 2e-6);
* mfs_factor_points: the KL factor with the Gram evaluated from the points,
  against the oracle's factor of the same scene, with the FP32 Gram's products
  on FP32 MFMA and as split BF16 products;
* mfs_solve_points: relative normal-equation residual <= tol, recomputed by the
  oracle on the FP64 kernel matrix; fixed-iteration trajectories against the
  oracle's PCG with the same factor values; warm start; error paths.
"""
import numpy as np
import pytest
import torch

import oracle
from fastmbcui_amd import _lib, scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _small(seed=0, k=2):
    return scenes.coins(seed=seed, n_bodies=3, pts_per_body=160, k=k, box=6.0)


def _ellips(seed=0, k=2):
    return scenes.ellipsoids(seed=seed, n_bodies=4, pts_per_body=256, k=k, box=8.0)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _points_ctx(sc, precision=_lib.MFS_FP32, rho=3.0):
    ctx = _lib.Context(0)
    ctx.plan(sc.y, rho)
    ctx.factor_points(_dev(sc.x), _dev(sc.nx), _dev(sc.y), precision)
    return ctx


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("k", [1, 3, 20])
def test_points_apply_fp64(trans, k):
    sc = _ellips(k=1)
    print(f"points apply fp64 trans={trans} k={k} n={sc.n}", flush=True)
    ctx = _points_ctx(sc)
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(((sc.m if trans else sc.n), k))
    got = ctx.points_apply(_dev(v), trans=trans, precision=_lib.MFS_FP64).cpu().numpy()
    ref = (M.T @ v) if trans else (M @ v)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"  max err / max |ref| = {err:.2e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("maker", [_small, _ellips])
def test_points_apply_fp32(trans, maker):
    sc = maker(k=1)
    print(f"points apply fp32 {sc.name} trans={trans}", flush=True)
    ctx = _points_ctx(sc)
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    rng = np.random.default_rng(6)
    v = rng.standard_normal(((sc.m if trans else sc.n), 8)).astype(np.float32)
    got = ctx.points_apply(_dev(v), trans=trans).cpu().numpy().astype(np.float64)
    A = M.T if trans else M
    ref = A @ v.astype(np.float64)
    # error relative to the absolute sum of the terms of each entry
    scale = np.abs(A) @ np.abs(v.astype(np.float64))
    err = np.max(np.abs(got - ref) / scale)
    print(f"  max entry err / sum|terms| = {err:.2e}", flush=True)
    assert err < 2e-6


def test_points_apply_near_pairs_keep_fp32_accuracy():
    # Sources 1e-3 inside targets far from the origin: a plain FP32 difference
    # of the coordinates (~1e3) would lose ~3 digits; the hi/lo split does not.
    rng = np.random.default_rng(2)
    n = 512
    base = rng.uniform(-1.0, 1.0, (n, 3)) + 1000.0
    nx = rng.standard_normal((n, 3))
    nx /= np.linalg.norm(nx, axis=1, keepdims=True)
    y = base - 1e-3 * nx
    sc = scenes.Scene("near", base, nx, y, np.zeros((n, 1)), np.zeros(n, np.int32), np.zeros((1, 3)))
    ctx = _points_ctx(sc)
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    v = rng.standard_normal((n, 1)).astype(np.float32)
    got = ctx.points_apply(_dev(v)).cpu().numpy()[:, 0].astype(np.float64)
    ref = (M @ v.astype(np.float64))[:, 0]
    scale = np.abs(M) @ np.abs(v[:, 0].astype(np.float64))
    err = np.max(np.abs(got - ref) / scale)
    print(f"near pairs: max entry err / sum|terms| = {err:.2e}", flush=True)
    assert err < 2e-6


def _col_rel_err(pl, a, b):
    worst = 0.0
    for j in range(pl.n):
        s, e = pl.sptr[j], pl.sptr[j + 1]
        worst = max(worst, np.linalg.norm(a[s:e] - b[s:e]) / np.linalg.norm(b[s:e]))
    return worst


@pytest.mark.parametrize("precision", [_lib.MFS_FP64, _lib.MFS_FP32])
def test_factor_points_close_to_oracle(precision):
    sc = _ellips()
    print(f"factor points precision={precision} n={sc.n}", flush=True)
    ctx = _points_ctx(sc, precision)
    val = ctx.factor_export()
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    pl = oracle.plan(sc.y, 3.0)
    ref = oracle.factor(oracle.permuted_colmajor(M, pl.Pinv), M.shape[0], pl)
    err = _col_rel_err(pl, val, ref)
    print(f"  max column rel err {err:.2e}, retried {ctx.stats().retried_supernodes}", flush=True)
    assert err < 1e-5


def test_factor_points_matches_dense_factor_on_same_plan():
    sc = _small()
    ctx = _points_ctx(sc, _lib.MFS_FP32)
    vp = ctx.factor_export()
    Md = ctx.assemble_dense(_dev(sc.x), _dev(sc.nx), _dev(sc.y))
    ctx.factor_dense(Md, _lib.MFS_FP32)
    vd = ctx.factor_export()
    pl = oracle.plan(sc.y, 3.0)
    err = _col_rel_err(pl, vp, vd)
    print(f"points vs dense fp32 factor: max column rel err {err:.2e}", flush=True)
    assert err < 5e-2


@pytest.mark.parametrize("maker,k", [(_small, 1), (_ellips, 3), (_ellips, 8)])
def test_solve_points_reaches_tolerance(maker, k):
    sc = maker(k=k)
    print(f"solve points {sc.name} k={k}", flush=True)
    ctx = _points_ctx(sc)
    B = _dev(sc.b.astype(np.float32))
    X = ctx.solve_points(B, maxit=500, tol=1e-6)
    res = ctx.points_residual(B, X)
    s = ctx.stats()
    print(f"  residual {res.max():.2e} outer {s.outer_iters} iters {[s.iters[c] for c in range(k)]}",
          flush=True)
    assert res.max() <= 1e-6
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    res_o = oracle.normal_residual(M, sc.b.astype(np.float32), X.cpu().numpy())
    assert res_o <= 2e-6


@pytest.mark.parametrize("k", [3, 8])
def test_solve_points_fixed_iterations_match_oracle_trajectory(k):
    """The matrix-free batched FP32 PCG after 1 and 3 iterations from zero
    (one sweep, no refinement) against the oracle's FP64 PCG on the FP64
    kernel matrix, run with the GPU's own factor values: the on-the-fly M p
    and M^T y passes, the preconditioner products and the updates.  Their
    FP32 evaluation moves x by 5.2e-7 .. 6.1e-7 (measured); a wrong target,
    source or sign moves it at O(1)."""
    sc = _ellips(k=k)
    ctx = _points_ctx(sc)
    val = ctx.factor_export()
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    pl = oracle.plan(sc.y, 3.0)
    Mp = oracle.permuted_colmajor(M, pl.Pinv)
    B = _dev(sc.b.astype(np.float32))
    for iters in (1, 3):
        X = ctx.solve_points(B, maxit=iters, tol=1e-30).cpu().numpy().astype(np.float64)
        s = ctx.stats()
        assert s.outer_iters == 1
        worst = 0.0
        for c in range(k):
            b = sc.b[:, c].astype(np.float32).astype(np.float64)
            xo, it, _ = oracle.pcg(Mp, M.shape[0], pl, val, b, 0, iters, 0.0)
            assert it == iters and s.iters[c] == iters, (c, it, s.iters[c])
            xo = xo[pl.P]
            worst = max(worst, np.linalg.norm(X[:, c] - xo) / np.linalg.norm(xo))
        print(f"points k={k} after {iters} iterations: max rel. diff {worst:.2e}", flush=True)
        assert worst < 6e-6, (k, iters, worst)


def test_solve_points_config1_coins():
    sc = scenes.coins(seed=0, k=1)   # N = 4096
    ctx = _points_ctx(sc)
    B = _dev(sc.b.astype(np.float32))
    X = ctx.solve_points(B, maxit=1000, tol=1e-6)
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    res_o = oracle.normal_residual(M, sc.b.astype(np.float32), X.cpu().numpy())
    print(f"coins N=4096 points: oracle residual {res_o:.2e} iters {ctx.stats().iters[0]}", flush=True)
    assert res_o <= 2e-6


def test_points_residual_matches_oracle():
    sc = _ellips(k=3)
    ctx = _points_ctx(sc)
    rng = np.random.default_rng(9)
    X = rng.standard_normal((sc.n, 3)).astype(np.float32) * 1e-2
    B = sc.b.astype(np.float32)
    got = ctx.points_residual(_dev(B), _dev(X))
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    for c in range(3):
        want = oracle.normal_residual(M, B[:, c:c + 1], X[:, c:c + 1])
        # numerator M^T(b - M x) in FP64; the denominator ||M^T b|| is summed
        # from FP32 kernel values (it only scales the stopping test)
        assert abs(got[c] - want) <= 1e-6 * want


def test_solve_points_warm_start_from_solution():
    sc = _ellips(k=2)
    ctx = _points_ctx(sc)
    B = _dev(sc.b.astype(np.float32))
    X = ctx.solve_points(B, tol=1e-6)
    cold = sum(ctx.stats().iters[c] for c in range(2))
    X2 = ctx.solve_points(B, tol=1e-6, x0=X)
    s = ctx.stats()
    warm = sum(s.iters[c] for c in range(2))
    print(f"warm start: cold {cold} inner iterations, warm {warm}, outer {s.outer_iters}", flush=True)
    assert s.outer_iters <= 1 and 3 * warm <= cold
    assert ctx.points_residual(B, X2).max() <= 1e-6


def test_points_entry_points_check_state():
    sc = _small(k=1)
    ctx = _lib.Context(0)
    ctx.plan(sc.y, 3.0)
    ctx.m_points = sc.m
    with pytest.raises(_lib.MfsError) as e:
        ctx.solve_points(_dev(sc.b.astype(np.float32)))
    assert e.value.code == _lib.MFS_ERR_STATE
    with pytest.raises(ValueError):
        ctx.factor_points(_dev(sc.x), _dev(sc.nx), _dev(sc.y[:-1]))


@pytest.mark.parametrize("precision", [_lib.MFS_FP64, _lib.MFS_FP32])
def test_points_large_supernodes(precision):
    sc = scenes.coins(seed=3, n_bodies=4, pts_per_body=512, k=2, box=6.0)
    ctx = _points_ctx(sc, precision, rho=7.0)
    val = ctx.factor_export()
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    pl = oracle.plan(sc.y, 7.0)
    ref = oracle.factor(oracle.permuted_colmajor(M, pl.Pinv), M.shape[0], pl)
    err = _col_rel_err(pl, val, ref)
    B = _dev(sc.b.astype(np.float32))
    X = ctx.solve_points(B, maxit=500, tol=1e-6)
    res = oracle.normal_residual(M, sc.b.astype(np.float32), X.cpu().numpy())
    print(f"points large supernodes (max local {ctx.stats().max_local}): factor err {err:.2e}, "
          f"residual {res:.2e}", flush=True)
    assert ctx.stats().max_local > 192
    assert res <= 2e-6


def _coins_large():
    return scenes.coins(seed=3, n_bodies=4, pts_per_body=512, k=2, box=6.0)


def _coins_ragged():
    return scenes.coins(seed=5, n_bodies=3, pts_per_body=150, k=2, box=6.0)   # m = 450: a partial 32-row step


@pytest.mark.parametrize("maker,rho,tol", [(_ellips, 3.0, 1e-5), (_coins_large, 7.0, None),
                                           (_coins_ragged, 3.0, None)])
def test_factor_points_gram_products(maker, rho, tol):
    """The FP32 points factor with its Gram's products on FP32 MFMA (1) and as
    three BF16 products of a hi + lo split of each evaluated value (2, the
    default; 8 rows per lane per 32-row step), against the oracle's FP64
    factor; rho 7 gives supernodes of 12 and 16 tiles; 450 targets end in a
    partial 32-row step (rows past the end read the last row with a zero
    scale: their values must be exactly zero).  The split leaves a
    relative error of about 2^-16 per product, so its factor may be a few
    times further from the FP64 one than the FP32 products' but within the
    FP32 factor bound; the solve on it still reaches its tolerance."""
    sc = maker()
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    pl = oracle.plan(sc.y, rho)
    ref = oracle.factor(oracle.permuted_colmajor(M, pl.Pinv), M.shape[0], pl)
    errs = {}
    for products in (1, 2):
        ctx = _lib.Context(0)
        ctx.set_param(_lib.MFS_PARAM_GRAM_SPLIT, products)
        ctx.plan(sc.y, rho)
        ctx.factor_points(_dev(sc.x), _dev(sc.nx), _dev(sc.y), _lib.MFS_FP32)
        errs[products] = _col_rel_err(pl, ctx.factor_export(), ref)
        st = ctx.stats()
        print(f"{sc.name} rho {rho}: Gram products {products}: max column rel err {errs[products]:.2e}, "
              f"max local {st.max_local}, retried {st.retried_supernodes}", flush=True)
        if products == 2:
            B = _dev(sc.b.astype(np.float32))
            X = ctx.solve_points(B, maxit=500, tol=1e-6)
            res = oracle.normal_residual(M, sc.b.astype(np.float32), X.cpu().numpy())
            assert res <= 2e-6
    if tol is not None:  # the bound of the FP32 points factor on this scene
        assert errs[1] < tol and errs[2] < tol
    assert errs[2] <= 20 * errs[1] + 1e-7


def test_time_stepper_warm_start_meets_tolerance_each_step():
    from fastmbcui_amd.timestep import TimeStepper
    rs = scenes.mixed_bodies(seed=1, n_bodies=6, pts_per_body=256, box=10.0)
    ts = TimeStepper(rs, rho=4.0, tol=1e-8, check_residual=True)
    infos = [ts.step()] + [ts.step(1e-2) for _ in range(3)]
    for s in infos:
        print(f"step {s.step}: iters {s.inner_iters} sweeps {s.sweeps} residual {s.residual:.2e} "
              f"warm {s.warm}", flush=True)
        assert s.residual <= 1e-8
    assert all(s.warm for s in infos[1:])
    assert max(s.inner_iters for s in infos[1:]) < infos[0].inner_iters


def test_autotune_points_blocks_keeps_results():
    from fastmbcui_amd import autotune
    sc = _ellips(k=2)
    ctx = _points_ctx(sc)
    B = _dev(sc.b.astype(np.float32))
    X0 = ctx.solve_points(B, tol=1e-6)
    r = autotune.tune_points_blocks(ctx, k=2, candidates=(512, 2048, 8192), reps=2)
    print(f"points blocks tuning: {r}", flush=True)
    assert r["best"] in (512, 2048, 8192)
    X1 = ctx.solve_points(B, tol=1e-6)
    assert ctx.points_residual(B, X1).max() <= 1e-6
    assert torch.allclose(X0, X1, rtol=1e-3, atol=1e-3 * X0.abs().max().item())
