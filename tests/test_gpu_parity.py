"""GPU parity of the HIP path against the CPU oracle, through the C-ABI.

* plan (ordering, pattern, supernodes, supernodal pattern): bit-exact;
* FP64 factor values: per-column relative error against the oracle;
* PCG: fixed-iteration trajectories and converged solutions;
* operators: MinvF_float_fastGrad / VpL_fast_float against the oracle's FP64
  restatement of the op, and the autograd formula against a dense FP64 solve.
"""
import numpy as np
import pytest
import torch

import oracle
from fastmbcui_amd import _lib, runtime, scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _small(seed=0, k=2):
    return scenes.coins(seed=seed, n_bodies=3, pts_per_body=160, k=k, box=6.0)


def _ellips(seed=0, k=2):
    return scenes.ellipsoids(seed=seed, n_bodies=4, pts_per_body=256, k=k, box=8.0)


def _M32(sc):
    return scenes.kernel_matrix(sc.x, sc.nx, sc.y).astype(np.float32)


def _ctx():
    return _lib.Context(0)


def _plan_pair(points64, rho=3.0, lam=1.5, cap=2048):
    ctx = _ctx()
    ctx.plan(points64, rho, lam, cap)
    got = ctx.plan_export()
    want = oracle.plan(points64, rho, lam, cap)
    return ctx, got, want


def _assert_plan_equal(got, want):
    for key in ("P", "Pinv", "scale", "ptr", "ind", "sup_ptr", "sup_ind", "sup_parent", "sptr", "sind"):
        g = got[key]
        w = getattr(want, key)
        assert g.shape == w.shape, key
        assert np.array_equal(g, w), key


# ----------------------------------------------------------------------------- plan
@pytest.mark.parametrize("maker", [_small, _ellips])
def test_plan_bit_exact(maker):
    sc = maker()
    print(f"plan parity: {sc.name} n={sc.n}", flush=True)
    _, got, want = _plan_pair(sc.y)
    _assert_plan_equal(got, want)


def test_plan_bit_exact_config1_coins():
    sc = scenes.coins(seed=0, k=1)   # N = 4096
    _, got, want = _plan_pair(sc.y)
    _assert_plan_equal(got, want)


def test_plan_from_float32_points_matches_oracle_on_upcast():
    sc = _small()
    y32 = sc.y.astype(np.float32)
    ctx = _ctx()
    ctx.plan(y32, 3.0)
    _assert_plan_equal(ctx.plan_export(), oracle.plan(y32.astype(np.float64), 3.0))


@pytest.mark.parametrize("rho,cap", [(1.2, 2048), (4.0, 2048), (3.0, 4)])
def test_plan_bit_exact_other_parameters(rho, cap):
    sc = _small(seed=2)
    _, got, want = _plan_pair(sc.y, rho, 1.5, cap)
    _assert_plan_equal(got, want)


def test_plan_random_cloud_and_tiny_sizes():
    rng = np.random.default_rng(11)
    for n in (2, 3, 7, 65, 300):
        pts = rng.uniform(-1, 1, (n, 3))
        _, got, want = _plan_pair(pts)
        _assert_plan_equal(got, want)


def test_plan_rejects_coincident_points():
    pts = np.random.default_rng(1).uniform(-1, 1, (40, 3))
    pts[7] = pts[30]
    ctx = _ctx()
    with pytest.raises(_lib.MfsError) as e:
        ctx.plan(pts, 3.0)
    assert e.value.code == _lib.MFS_ERR_DUPLICATE


# ----------------------------------------------------------------------------- factor
def _factor_pair(sc, precision=_lib.MFS_FP64, rho=3.0, gram_split=0):
    M = _M32(sc)
    ctx, got, want = _plan_pair(sc.y, rho)
    Md = torch.from_numpy(M).to(DEV)
    ctx.set_param(_lib.MFS_PARAM_GRAM_SPLIT, gram_split)
    ctx.factor_dense(Md, precision)
    val = ctx.factor_export()
    Mp = oracle.permuted_colmajor(M.astype(np.float64), want.Pinv)
    ref = oracle.factor(Mp, M.shape[0], want)
    return ctx, Md, M, want, val, ref


def _col_rel_err(pl, a, b):
    worst = 0.0
    for j in range(pl.n):
        s, e = pl.sptr[j], pl.sptr[j + 1]
        worst = max(worst, np.linalg.norm(a[s:e] - b[s:e]) / np.linalg.norm(b[s:e]))
    return worst


@pytest.mark.parametrize("maker", [_small, _ellips])
def test_factor_fp64_matches_oracle(maker):
    sc = maker()
    ctx, Md, M, pl, val, ref = _factor_pair(sc)
    err = _col_rel_err(pl, val, ref)
    print(f"factor fp64 {sc.name}: max column rel err {err:.2e}", flush=True)
    assert err < 1e-9


@pytest.mark.parametrize("split", [1, 2])
def test_factor_fp32_close_to_oracle(split):
    """FP32 factor with the Gram's products on FP32 MFMA (1) or as three BF16
    products of a hi + lo split of each value (2), against the oracle's FP64
    factor: 1.0e-6 and 2.7e-6 measured.  The split's dropped terms are below
    about 3 2^-16 of each product; dropping one of the two cross products
    instead (2^-9 of a product) moves the Gram by ~1e-3 and fails the bound."""
    sc = _ellips()
    ctx, Md, M, pl, val, ref = _factor_pair(sc, _lib.MFS_FP32, gram_split=split)
    err = _col_rel_err(pl, val, ref)
    print(f"factor fp32 (Gram products {split}): max column rel err {err:.2e}, "
          f"retried {ctx.stats().retried_supernodes}", flush=True)
    assert err < 5e-5


@pytest.mark.parametrize("ppb", [24, 50, 64])
def test_factor_small_m_matches_oracle(ppb):
    """m <= 128 (M^T rows padded to 128 only): the Gram's row slabs stay within
    a row for every wave of the small-tile path; FP64 and FP32 factors (both
    kinds of FP32 Gram products)."""
    sc = scenes.coins(seed=4, n_bodies=2, pts_per_body=ppb, k=1, box=4.0)
    assert sc.m <= 128
    for precision, tol, split in ((_lib.MFS_FP64, 1e-9, 0), (_lib.MFS_FP32, 5e-2, 1),
                                  (_lib.MFS_FP32, 5e-2, 2)):
        ctx, Md, M, pl, val, ref = _factor_pair(sc, precision, gram_split=split)
        assert _col_rel_err(pl, val, ref) < tol, (precision, split)


def test_factor_dense_pattern_one_iteration():
    sc = scenes.coins(seed=1, n_bodies=2, pts_per_body=24, k=1, box=4.0)
    M = _M32(sc)
    ctx = _ctx()
    ctx.plan(sc.y, 1e6)
    Md = torch.from_numpy(M).to(DEV)
    ctx.factor_dense(Md)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    ctx.solve_dense(Md, B, maxit=10, tol=1e-8)
    assert ctx.stats().iters[0] == 1


# ----------------------------------------------------------------------------- PCG
def test_pcg_fixed_iterations_match_oracle_trajectory():
    sc = _small(k=2)
    ctx, Md, M, pl, val, ref = _factor_pair(sc)
    Mp = oracle.permuted_colmajor(M.astype(np.float64), pl.Pinv)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    for iters in (1, 3, 6):
        X = ctx.solve_dense(Md, B, maxit=iters, tol=0.0).cpu().numpy()
        for c in range(sc.k):
            b = sc.b[:, c].astype(np.float32).astype(np.float64)
            xo, it, _ = oracle.pcg(Mp, M.shape[0], pl, ref, b, 0, iters, 0.0)
            assert it == iters and ctx.stats().iters[c] == iters
            xo = xo[pl.P]
            rel = np.linalg.norm(X[:, c] - xo) / np.linalg.norm(xo)
            assert rel < 1e-6, (iters, c, rel)   # FP32 output rounding dominates


@pytest.mark.parametrize("ppb", [161, 162, 163])
def test_pcg_ragged_sizes_match_oracle_trajectory(ppb):
    """n % 4 = 3, 2, 1: the matvec passes' scalar paths (rows not 16-byte
    aligned) and their small-grid launch shapes, against the oracle."""
    sc = scenes.coins(seed=2, n_bodies=3, pts_per_body=ppb, k=2, box=6.0)
    assert sc.n % 4 != 0
    ctx, Md, M, pl, val, ref = _factor_pair(sc)
    Mp = oracle.permuted_colmajor(M.astype(np.float64), pl.Pinv)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    for iters in (1, 4):
        X = ctx.solve_dense(Md, B, maxit=iters, tol=0.0).cpu().numpy()
        for c in range(sc.k):
            b = sc.b[:, c].astype(np.float32).astype(np.float64)
            xo, it, _ = oracle.pcg(Mp, M.shape[0], pl, ref, b, 0, iters, 0.0)
            xo = xo[pl.P]
            rel = np.linalg.norm(X[:, c] - xo) / np.linalg.norm(xo)
            assert rel < 1e-6, (ppb, iters, c, rel)
    ctx.factor_dense(Md, _lib.MFS_FP32)
    Xm = ctx.solve_mixed(Md, B, maxit=500, tol=1e-6).cpu().numpy().astype(np.float64)
    res = oracle.normal_residual(M.astype(np.float64), sc.b.astype(np.float32), Xm)
    assert res < 1e-6, res


@pytest.mark.parametrize("k", [8, 13])
def test_mixed_fixed_iterations_match_oracle_trajectory(k):
    """The batched FP32 PCG of the chosen form (batch of 8; 13 padded to 16):
    its preconditioner products L^T r and L t (+ dot), vector updates and
    alpha/beta, against the oracle's FP64 PCG run with the GPU's own FP32
    factor values, after 1 and 3 iterations from zero (one sweep, no
    refinement).  Rows of L at rho = 3 hold from 1 to over 100 entries, so
    the products' four-entries-at-a-time steps and their one-entry remainders
    both run; a wrong index or a dropped entry moves x by far more than the
    tolerance, FP32 arithmetic by 3.8e-7 .. 4.6e-7 (measured)."""
    sc = _ellips(k=k)
    ctx, Md, M, pl, val, ref = _factor_pair(sc, _lib.MFS_FP32)
    Mp = oracle.permuted_colmajor(M.astype(np.float64), pl.Pinv)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    for iters in (1, 3):
        X = ctx.solve_mixed(Md, B, maxit=iters, tol=1e-30).cpu().numpy().astype(np.float64)
        s = ctx.stats()
        assert s.outer_iters == 1
        worst = 0.0
        for c in range(k):
            b = sc.b[:, c].astype(np.float32).astype(np.float64)
            xo, it, _ = oracle.pcg(Mp, M.shape[0], pl, val, b, 0, iters, 0.0)
            assert it == iters and s.iters[c] == iters, (c, it, s.iters[c])
            xo = xo[pl.P]
            worst = max(worst, np.linalg.norm(X[:, c] - xo) / np.linalg.norm(xo))
        print(f"mixed k={k} after {iters} iterations: max rel. diff {worst:.2e}", flush=True)
        assert worst < 5e-6, (k, iters, worst)


def test_pcg_converged_matches_direct_solve():
    sc = _small(k=2)
    ctx, Md, M, pl, val, ref = _factor_pair(sc)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    X = ctx.solve_dense(Md, B, maxit=500, tol=1e-11).cpu().numpy().astype(np.float64)
    M64 = M.astype(np.float64)
    xd = np.linalg.solve(M64.T @ M64, M64.T @ sc.b.astype(np.float32).astype(np.float64))
    assert np.linalg.norm(X - xd) / np.linalg.norm(xd) < 1e-6
    assert oracle.normal_residual(M64, sc.b.astype(np.float32), X) < 1e-6


def test_pcg_iteration_counts_match_oracle_at_reference_tolerance():
    sc = _ellips(k=3)
    ctx, Md, M, pl, val, ref = _factor_pair(sc)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    X = ctx.solve_dense(Md, B, maxit=200, tol=1e-6).cpu().numpy()
    r = oracle.minvf(M.astype(np.float64), sc.b.astype(np.float32), sc.y, 3.0, 200, 1e-6, pl=pl, val=ref)
    got = [ctx.stats().iters[c] for c in range(sc.k)]
    assert all(abs(a - b) <= 1 for a, b in zip(got, r.iters)), (got, r.iters)
    for c in range(sc.k):
        if got[c] == r.iters[c]:
            rel = np.linalg.norm(X[:, c] - r.x[:, c]) / np.linalg.norm(r.x[:, c])
            assert rel < 1e-5


def test_pcg_zero_rhs_and_zero_iterations():
    sc = _small(k=2)
    ctx, Md, M, pl, val, ref = _factor_pair(sc)
    B = torch.zeros((M.shape[0], 2), dtype=torch.float32, device=DEV)
    X = ctx.solve_dense(Md, B, maxit=50, tol=1e-6)
    assert torch.count_nonzero(X).item() == 0 and ctx.stats().iters[0] == 0
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    X = ctx.solve_dense(Md, B, maxit=0, tol=1e-6)
    assert torch.count_nonzero(X).item() == 0


def test_source_space_solve_matches_oracle_vpl():
    sc = _small(k=1)
    ctx, Md, M, pl, val, ref = _factor_pair(sc)
    rng = np.random.default_rng(5)
    G = rng.standard_normal((sc.n, 2)).astype(np.float32)
    X = ctx.solve_dense(Md, torch.from_numpy(G).to(DEV), rhs_space=_lib.MFS_RHS_SOURCE,
                        maxit=500, tol=1e-11).cpu().numpy()
    M64 = M.astype(np.float64)
    xd = np.linalg.solve(M64.T @ M64, G.astype(np.float64))
    assert np.linalg.norm(X - xd) / np.linalg.norm(xd) < 1e-6


@pytest.mark.parametrize("k", [3, 18])
def test_mixed_source_space_solve_reaches_tolerance(k):
    """The batched mixed solve of the VpL system M^T M x = b: FP64 residual
    ||b - M^T M x|| <= tol ||b|| per column (recomputed here), two batches for
    k = 18."""
    sc = _small(k=1)
    ctx, Md, M, pl, val, ref = _factor_pair(sc, _lib.MFS_FP32)
    rng = np.random.default_rng(k)
    G = rng.standard_normal((sc.n, k)).astype(np.float32)
    X = ctx.solve_mixed(Md, torch.from_numpy(G).to(DEV), maxit=2000, tol=1e-6,
                        rhs_space=_lib.MFS_RHS_SOURCE).cpu().numpy().astype(np.float64)
    M64 = M.astype(np.float64)
    G64 = G.astype(np.float64)
    A = M64.T @ M64
    res = lambda Y: np.linalg.norm(G64 - A @ Y, axis=0) / np.linalg.norm(G64, axis=0)
    # the solver certifies its FP64 iterate to 1e-6; the FP32 output adds the
    # residual that rounding the exact solution to FP32 already has
    floor = res(np.linalg.solve(A, G64).astype(np.float32).astype(np.float64))
    rel = res(X)
    assert np.all(rel <= 1e-6 + 2 * floor), (rel.max(), floor.max())
    s = ctx.stats()
    assert all(s.relres[c] <= 1e-6 for c in range(min(k, 64)))
    with pytest.raises(ValueError):
        ctx.solve_mixed(Md, torch.zeros((sc.m + 1, 2), device=DEV), rhs_space=_lib.MFS_RHS_SOURCE)


def test_op_autograd_mixed_mode_matches_plain():
    """MinvF_float_fastGrad's forward and backward (VpL) in the mixed mode agree
    with the plain mode to the solver tolerance."""
    from fastmbcui_amd import runtime
    from fastmbcui_amd.mfs_torch_py import ops
    sc = _small(k=3)
    M = torch.from_numpy(_M32(sc)).to(DEV).requires_grad_()
    F = torch.from_numpy(sc.b.astype(np.float32)).to(DEV).requires_grad_()
    Vs = torch.from_numpy(sc.y.astype(np.float32))
    W = torch.randn((sc.n, sc.k), device=DEV, generator=torch.Generator(DEV).manual_seed(3))
    out = {}
    try:
        for mode in ("plain", "mixed"):
            runtime.reset()
            runtime.set_solver_mode(mode)
            X = ops.MinvF_float_fastGrad(M, F, Vs, 3.0, 2000, 1e-8)
            gM, gF = torch.autograd.grad((X * W).sum(), (M, F))
            out[mode] = (X.detach(), gM, gF)
    finally:
        runtime.set_solver_mode("plain")
        runtime.reset()
    for a, b in zip(out["mixed"], out["plain"]):
        assert ((a - b).norm() / b.norm()).item() < 1e-4
@pytest.mark.parametrize("maker,k", [(_small, 2), (_ellips, 5)])
def test_mixed_solve_reaches_tolerance(maker, k):
    sc = maker(k=k)
    ctx, Md, M, pl, val, ref = _factor_pair(sc, _lib.MFS_FP32)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    X = ctx.solve_mixed(Md, B, maxit=500, tol=1e-6)
    res = ctx.normal_residual(Md, B, X)
    s = ctx.stats()
    print(f"mixed {sc.name}: residual {res.max():.2e} outer {s.outer_iters} "
          f"iters {[s.iters[c] for c in range(k)]}", flush=True)
    assert res.max() <= 1e-6
    res_o = oracle.normal_residual(M.astype(np.float64), sc.b.astype(np.float32), X.cpu().numpy())
    assert res_o <= 2e-6


# ----------------------------------------------------------------------------- operators
def test_minvf_op_matches_oracle():
    import fastmbcui_amd.mfs_torch_py as mfs
    runtime.reset()
    sc = _ellips(k=4)
    M = _M32(sc)
    Vs = torch.from_numpy(sc.y.astype(np.float32))
    F = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    out = mfs.ops.MinvF_float_fastGrad(torch.from_numpy(M).to(DEV), F, Vs, 3.0, 75, 1e-6)
    r = oracle.minvf(M.astype(np.float64), sc.b.astype(np.float32), Vs.numpy().astype(np.float64),
                     3.0, 75, 1e-6)
    info = mfs.last_solve_info()
    assert out.shape == (sc.n, 4) and out.dtype == torch.float32
    o = out.cpu().numpy()
    for c in range(4):
        assert abs(info.iters[c] - r.iters[c]) <= 1
        if info.iters[c] == r.iters[c]:
            assert np.linalg.norm(o[:, c] - r.x[:, c]) / np.linalg.norm(r.x[:, c]) < 1e-5


def test_vpl_requires_a_forward_first():
    import fastmbcui_amd.mfs_torch_py as mfs
    runtime.reset()
    sc = _small(k=1)
    M = torch.from_numpy(_M32(sc)).to(DEV)
    g = torch.ones((sc.n, 1), device=DEV)
    with pytest.raises(RuntimeError, match="Please run MinvF_float first"):
        mfs.ops.VpL_fast_float(M, g, torch.from_numpy(sc.y.astype(np.float32)), 3.0, 50, 1e-6)


def test_autograd_matches_dense_closed_form():
    import fastmbcui_amd.mfs_torch_py as mfs
    runtime.reset()
    sc = _small(k=2)
    M32 = _M32(sc)
    F32 = sc.b.astype(np.float32)
    Vs = torch.from_numpy(sc.y.astype(np.float32))
    M = torch.from_numpy(M32).to(DEV).requires_grad_(True)
    F = torch.from_numpy(F32).to(DEV).requires_grad_(True)
    rng = np.random.default_rng(9)
    G = rng.standard_normal((sc.n, 2)).astype(np.float32)
    A = mfs.ops.MinvF_float_fastGrad(M, F, Vs, 3.0, 500, 1e-11)
    (A * torch.from_numpy(G).to(DEV)).sum().backward()
    M64, F64, G64 = M32.astype(np.float64), F32.astype(np.float64), G.astype(np.float64)
    N = M64.T @ M64
    A64 = np.linalg.solve(N, M64.T @ F64)
    V = np.linalg.solve(N, G64)
    gF = M64 @ V
    gM = F64 @ V.T - M64 @ (A64 @ V.T + V @ A64.T)
    assert np.linalg.norm(F.grad.cpu().numpy() - gF) / np.linalg.norm(gF) < 1e-4
    assert np.linalg.norm(M.grad.cpu().numpy() - gM) / np.linalg.norm(gM) < 1e-4


def test_op_reuses_plan_and_factor_for_identical_inputs():
    import fastmbcui_amd.mfs_torch_py as mfs
    runtime.reset()
    sc = _small(k=1)
    M = torch.from_numpy(_M32(sc)).to(DEV)
    F = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    Vs = torch.from_numpy(sc.y.astype(np.float32))
    a = mfs.ops.MinvF_float_fastGrad(M, F, Vs, 3.0, 75, 1e-6)
    assert not mfs.last_solve_info().plan_reused
    b = mfs.ops.MinvF_float_fastGrad(M, F, Vs, 3.0, 75, 1e-6)
    info = mfs.last_solve_info()
    assert info.plan_reused and info.factor_reused
    assert torch.equal(a, b)
    M.mul_(1.0)   # in-place write bumps the version: factor must be rebuilt
    mfs.ops.MinvF_float_fastGrad(M, F, Vs, 3.0, 75, 1e-6)
    assert not mfs.last_solve_info().factor_reused


def test_op_factor_cache_sees_new_contents_at_a_recycled_address():
    """A new M of the same shape, placed by torch's caching allocator where a
    freed M was (fresh version counter), must not reuse the old factor."""
    import fastmbcui_amd.mfs_torch_py as mfs
    runtime.reset()
    sc = _small(k=1)
    M32 = _M32(sc)
    F = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    Vs = torch.from_numpy(sc.y.astype(np.float32))
    M1 = torch.from_numpy(M32).to(DEV)
    addr = M1.data_ptr()
    mfs.ops.MinvF_float_fastGrad(M1, F, Vs, 3.0, 75, 1e-6)
    del M1
    M2 = torch.from_numpy(M32 * np.float32(1.5)).to(DEV)
    print(f"recycled address: {M2.data_ptr() == addr}", flush=True)
    out = mfs.ops.MinvF_float_fastGrad(M2, F, Vs, 3.0, 500, 1e-10)
    assert not mfs.last_solve_info().factor_reused
    r = oracle.minvf((M32 * np.float32(1.5)).astype(np.float64), sc.b.astype(np.float32),
                     sc.y, 3.0, 500, 1e-10)
    assert np.linalg.norm(out.cpu().numpy() - r.x) / np.linalg.norm(r.x) < 1e-5


def test_assemble_tall_matrix_beyond_the_grid_row_limit():
    """m > 65,535 rows (more than one grid row per target would allow)."""
    rng = np.random.default_rng(4)
    m, n = 70001, 5
    x = rng.uniform(-1, 1, (m, 3))
    nx = rng.standard_normal((m, 3))
    nx /= np.linalg.norm(nx, axis=1, keepdims=True)
    y = rng.uniform(-1, 1, (n, 3)) * 0.3 + 3.0
    ctx = _ctx()
    M = ctx.assemble_dense(torch.from_numpy(x).to(DEV), torch.from_numpy(nx).to(DEV),
                           torch.from_numpy(y).to(DEV)).cpu().numpy()
    ref = scenes.kernel_matrix(x, nx, y)
    # FP32 rounding per entry; an absolute floor for entries where (x - y).n cancels
    assert np.all(np.abs(M - ref) <= 1e-6 * np.abs(ref) + 1e-12 * np.abs(ref).max())


def test_op_prints_the_reference_phase_timers(capsys):
    """timing_verbose prints the reference's eleven TIME(...) phases
    (mfs_torch_ops.cu:48-193), in its order."""
    import fastmbcui_amd.mfs_torch_py as mfs
    runtime.reset()
    sc = _small(k=1)
    M = torch.from_numpy(_M32(sc)).to(DEV)
    F = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    Vs = torch.from_numpy(sc.y.astype(np.float32))
    mfs.ops.MinvF_float_fastGrad(M, F, Vs, 3.0, 75, 1e-6, True, True)
    out = capsys.readouterr().out.splitlines()
    phases = [ln.split("(")[1].split(")")[0] for ln in out if ln.startswith("TIME(")]
    assert phases == ["prep", "fps_sampler", "permutation", "build_cov_mat", "pattern", "setup",
                      "compute", "cache", "pcg", "perm_back", "total"]
    assert any(ln.startswith("pcg column 0: iterations") for ln in out)


def test_op_rejects_coincident_sources():
    import fastmbcui_amd.mfs_torch_py as mfs
    runtime.reset()
    sc = _small(k=1)
    y = sc.y.astype(np.float32).copy()
    y[5] = y[17]
    M = torch.from_numpy(_M32(sc)).to(DEV)
    F = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    with pytest.raises(RuntimeError):
        mfs.ops.MinvF_float_fastGrad(M, F, torch.from_numpy(y), 3.0, 75, 1e-6)


def test_assemble_matches_numpy_kernel():
    sc = _small(k=1)
    ctx = _ctx()
    x = torch.from_numpy(sc.x).to(DEV)
    nx = torch.from_numpy(sc.nx).to(DEV)
    y = torch.from_numpy(sc.y).to(DEV)
    M = ctx.assemble_dense(x, nx, y).cpu().numpy()
    ref = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    assert np.max(np.abs(M - ref) / np.abs(ref)) < 1e-6


@pytest.mark.parametrize("rho", [3.0, 6.0])
def test_plan_grid_search_bit_exact_larger_scene(rho):
    # 32 coins (N = 16,384): the grid search covers most columns, the coarsest
    # bands use the exhaustive scan.
    sc = scenes.coins(seed=3, n_bodies=32, pts_per_body=512, k=1, box=16.0)
    ctx, got, want = _plan_pair(sc.y, rho)
    s = ctx.stats()
    print(f"rho {rho}: grid columns {s.plan_grid_columns}/{sc.n}, "
          f"max column {np.diff(got['ptr']).max()}, order {s.plan_order_ms:.1f} ms, "
          f"pattern {s.plan_pattern_ms:.1f} ms", flush=True)
    assert s.plan_grid_columns > 0
    _assert_plan_equal(got, want)


def _large_sc(k=2):
    # rho = 7 on 4 coins gives supernodes of 129..236 local rows (TH in global scratch)
    return scenes.coins(seed=3, n_bodies=4, pts_per_body=512, k=k, box=6.0)


def test_factor_large_supernodes_matches_oracle():
    sc = _large_sc()
    ctx, Md, M, pl, val, ref = _factor_pair(sc, _lib.MFS_FP64, rho=7.0)
    s = ctx.stats()
    err = _col_rel_err(pl, val, ref)
    print(f"large supernodes (max local {s.max_local}) FP64: max column rel err {err:.2e}", flush=True)
    assert s.max_local > 192
    # rho = 7 blocks of ~236 rows are far worse conditioned than the rho = 3
    # ones, so two FP64 factorizations summing in different orders agree to
    # ~1e-8 here rather than ~1e-13
    assert err < 1e-7


def test_large_supernodes_fp32_factor_still_preconditions():
    # At rho = 7 the FP32 Cholesky of the ~236-row blocks carries O(1e-1)
    # relative error (condition numbers ~1e6), so what is checked is its use:
    # the mixed solve reaches the tolerance in about the FP64 factor's count.
    sc = _large_sc()
    ctx, Md, M, pl, val, ref = _factor_pair(sc, _lib.MFS_FP64, rho=7.0)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    ctx.solve_dense(Md, B, maxit=500, tol=1e-6)
    it64 = max(ctx.stats().iters[c] for c in range(sc.k))
    ctx.factor_dense(Md, _lib.MFS_FP32)
    X = ctx.solve_mixed(Md, B, maxit=500, tol=1e-6)
    s = ctx.stats()
    it32 = max(s.iters[c] for c in range(sc.k))
    res = oracle.normal_residual(M.astype(np.float64), sc.b.astype(np.float32), X.cpu().numpy())
    print(f"large supernodes: FP64 factor {it64} iterations, FP32 factor {it32} inner iterations "
          f"({s.outer_iters} sweeps), residual {res:.2e}", flush=True)
    assert res <= 2e-6
    assert it32 <= 2 * it64 + 2


def test_supernode_above_256_rows_is_rejected():
    sc = scenes.ellipsoids(seed=0, n_bodies=4, pts_per_body=512, k=1, box=8.0)
    ctx = _ctx()
    with pytest.raises(_lib.MfsError) as e:
        ctx.plan(sc.y, 8.0)
    assert e.value.code == _lib.MFS_ERR_ARG and "256" in str(e.value)


@pytest.mark.parametrize("m,n,k", [(700, 512, 3), (1301, 1027, 8), (517, 300, 11), (64, 40, 16)])
def test_dense_grad_matches_fp64_formula(m, n, k):
    """mfs_dense_grad (one pass over M for M A and M V, one rank-2k write of
    grad_M) against the backward formula in FP64: grad_F = M V,
    grad_M = F V^T - M (A V^T + V A^T) (core/mfs_torch_py/ops.py:42-58);
    2k <= 16 and 2k > 16 (two row passes), ragged n."""
    g = torch.Generator().manual_seed(m + n + k)
    M = torch.randn(m, n, generator=g)
    F = torch.randn(m, k, generator=g)
    A = torch.randn(n, k, generator=g)
    V = torch.randn(n, k, generator=g)
    ctx = _ctx()
    gM, gF = ctx.dense_grad(M.to(DEV), F.to(DEV), A.to(DEV), V.to(DEV))
    M64, F64, A64, V64 = (t.double() for t in (M, F, A, V))
    W = A64 @ V64.T
    gM_ref = F64 @ V64.T - M64 @ (W + W.T)
    gF_ref = M64 @ V64
    assert ((gF.cpu().double() - gF_ref).norm() / gF_ref.norm()).item() < 1e-6
    assert ((gM.cpu().double() - gM_ref).norm() / gM_ref.norm()).item() < 1e-5
    gM2, gF2 = ctx.dense_grad(M.to(DEV), F.to(DEV), A.to(DEV), V.to(DEV), need_m=False)
    assert gM2 is None and torch.equal(gF2, gF)
