"""The one-pass normal matvec (k_fused.hip) through the C-ABI.

* Q = M^T (M X) on small-integer data, where every FP32 partial sum is exact:
  bit-equal to the integer product through both passes and the one-pass kernel;
* Q = M^T (M X) from mfs_dense_normal_apply against the same product in FP64
  (numpy / torch on FP64 copies of the FP32 inputs, whose products are
  exact): the one-pass kernel and the two-pass matvec both to FP32-level
  relative error, at ragged row panels (m not a multiple of 16), a partial
  last column slice, a batch of 5 padded to 8 and the full 65,536 x 65,536
  workload (64 slices of 1024 columns x 4 row groups, one block per CU);
* the mixed solve, whose answer is certified by the FP64 two-pass residual
  M^T (b - M x), reaches its tolerance with the one-pass matvec, boundary and
  source-space right-hand sides, and the batch (which runs until its slowest
  column converges) takes no more iterations than with two passes, up to one
  (per column the counts scatter both ways with the FP32 rounding);
* shapes the kernel does not take (odd n, batches wider than 8 or of 4
  and fewer, n above 65,536) use two passes, and so does the default mode
  below 2^27 entries of M.
"""
import numpy as np
import pytest
import torch

from fastmbcui_amd import _lib, scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)
DEFAULT, OFF, FORCE = 0, 1, 2


def _solve(ctx, Md, B, mode, rhs_space=None, tol=1e-6):
    ctx.set_param(_lib.MFS_PARAM_FUSED, mode)
    ctx.track([_lib.MFS_K_MATVEC_FUSED])
    X = ctx.solve_mixed(Md, B, maxit=2000, tol=tol, rhs_space=rhs_space)
    torch.cuda.synchronize()
    launches = ctx.tracked(_lib.MFS_K_MATVEC_FUSED)[0]
    ctx.track([])
    s = ctx.stats()
    iters = [int(s.iters[c]) for c in range(B.shape[1])]
    ctx.set_param(_lib.MFS_PARAM_FUSED, DEFAULT)
    return X, launches, iters


def _setup(sc, extra_rows=0, seed=0):
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y).astype(np.float32)
    B = sc.b.astype(np.float32)
    if extra_rows:
        rng = np.random.default_rng(seed)
        idx = rng.choice(M.shape[0], extra_rows, replace=False)
        M = np.vstack([M, 0.5 * M[idx]])
        B = np.vstack([B, rng.standard_normal((extra_rows, B.shape[1])).astype(np.float32)])
    ctx = _lib.Context(0)
    ctx.plan(sc.y, 3.0)
    Md = torch.from_numpy(M).to(DEV)
    ctx.factor_dense(Md, _lib.MFS_FP32)
    return ctx, Md, torch.from_numpy(B).to(DEV)


def _compare(ctx, Md, B, want_fused=True, rhs_space=None):
    Xf, lf, itf = _solve(ctx, Md, B, FORCE, rhs_space)
    X2, l2, it2 = _solve(ctx, Md, B, OFF, rhs_space)
    print(f"fused: m={Md.shape[0]} n={Md.shape[1]} k={B.shape[1]} launches {lf}, iters {itf} vs {it2}",
          flush=True)
    assert l2 == 0
    assert (lf > 0) == want_fused
    if rhs_space is None:
        res = ctx.normal_residual(Md, B, Xf)
        assert res.max() <= 1e-6, res
    assert max(itf) <= max(it2) + 1, (itf, it2)
    return Xf


def _apply(ctx, Md, X, mode):
    ctx.set_param(_lib.MFS_PARAM_FUSED, mode)
    ctx.track([_lib.MFS_K_MATVEC_FUSED])
    Q = ctx.dense_normal_apply(Md, X)
    launches = ctx.tracked(_lib.MFS_K_MATVEC_FUSED)[0]
    ctx.track([])
    ctx.set_param(_lib.MFS_PARAM_FUSED, DEFAULT)
    return Q, launches


def _ref_normal(Md, X, chunk=8192):
    """M^T (M X) in FP64 on the device, in row chunks of M."""
    X64 = X.double()
    Q = torch.zeros((Md.shape[1], X.shape[1]), dtype=torch.float64, device=Md.device)
    for r0 in range(0, Md.shape[0], chunk):
        Mc = Md[r0:r0 + chunk].double()
        Q += Mc.t() @ (Mc @ X64)
    return Q


def _check_apply(Md, X, want_fused=True, tol=2e-6):
    ctx = _lib.Context(0)
    ref = _ref_normal(Md, X)
    for mode in (FORCE, OFF):
        Q, launches = _apply(ctx, Md, X, mode)
        assert (launches > 0) == (want_fused and mode == FORCE), (mode, launches)
        err = ((Q - ref).norm(dim=0) / ref.norm(dim=0)).max().item()
        print(f"normal apply m={Md.shape[0]} n={Md.shape[1]} k={X.shape[1]} mode {mode}: "
              f"one-pass launches {launches}, max rel. error {err:.2e}", flush=True)
        assert err < tol, (mode, err)


@pytest.mark.parametrize("m,bodies,per_body,k", [(16411, 4, 501, 8), (16411, 3, 667, 5)])
def test_normal_apply_exact_on_small_integers(m, bodies, per_body, k):
    """M, X with entries in -2..2: every FP32 product and partial sum of both
    passes (rows of M X: |.| <= 4 n; 32-row chunks of M^T y: <= 256 n) is an
    integer below 2^24, and the chunk sums are carried in FP64, so Q must equal
    the integer product exactly -- the production passes at m >= 16,384 (n = 2004
    and n = 2001, not a multiple of 4) and the one-pass kernel where it applies."""
    sc = scenes.coins(seed=1, n_bodies=bodies, pts_per_body=per_body, k=1, box=8.0)
    n = sc.n
    rng = np.random.default_rng(m + n + k)
    M = rng.integers(-2, 3, (m, n)).astype(np.float32)
    X = rng.integers(-2, 3, (n, k)).astype(np.float32)
    ref = M.astype(np.int64).T @ (M.astype(np.int64) @ X.astype(np.int64))
    ctx = _lib.Context(0)
    Md = torch.from_numpy(M).to(DEV)
    Xd = torch.from_numpy(X).to(DEV)
    for mode in (FORCE, OFF):
        Q, launches = _apply(ctx, Md, Xd, mode)
        got = Q.cpu().numpy()
        print(f"exact normal apply m={m} n={n} k={k} mode {mode}: one-pass launches {launches}, "
              f"max abs diff {np.abs(got - ref).max():.3g}", flush=True)
        assert np.array_equal(got, ref.astype(np.float64)), (mode, np.abs(got - ref).max())


@pytest.mark.parametrize("m,n,k", [(1, 4, 8), (5, 8, 8), (70, 12, 3), (129, 1028, 8), (40, 4096, 8),
                                   (300, 2052, 8)])
def test_normal_apply_exact_tiny_shapes(m, n, k):
    """Degenerate shapes through both matvecs: a single row, fewer rows than a
    panel or a 64-row chunk, ragged last panels and chunks, partial column
    slices and blocks, row groups of one or two panels (the one-pass kernel's
    last panels, whose reduce cannot wait for a later step); small integers,
    so Q must equal the integer product exactly."""
    rng = np.random.default_rng(m * 1000 + n + k)
    M = rng.integers(-2, 3, (m, n)).astype(np.float32)
    X = rng.integers(-2, 3, (n, k)).astype(np.float32)
    ref = M.astype(np.int64).T @ (M.astype(np.int64) @ X.astype(np.int64))
    ctx = _lib.Context(0)
    Md, Xd = torch.from_numpy(M).to(DEV), torch.from_numpy(X).to(DEV)
    eligible = k == 8 and n % 2 == 0
    for mode in (OFF, FORCE, DEFAULT):      # the default keeps two passes at these sizes
        Q, launches = _apply(ctx, Md, Xd, mode)
        assert (launches > 0) == (mode == FORCE and eligible), (mode, launches)
        got = Q.cpu().numpy()
        assert np.array_equal(got, ref.astype(np.float64)), (mode, np.abs(got - ref).max())


def test_normal_apply_wider_than_the_one_pass_grid():
    """More than 64 slices of 1024 columns (n > 65,536): two passes."""
    rng = np.random.default_rng(11)
    m, n = 40, 65540
    M = rng.integers(-2, 3, (m, n)).astype(np.float32)
    X = rng.integers(-2, 3, (n, 8)).astype(np.float32)
    ref = M.astype(np.int64).T @ (M.astype(np.int64) @ X.astype(np.int64))
    ctx = _lib.Context(0)
    Q, launches = _apply(ctx, torch.from_numpy(M).to(DEV), torch.from_numpy(X).to(DEV), FORCE)
    assert launches == 0
    assert np.array_equal(Q.cpu().numpy(), ref.astype(np.float64))


@pytest.mark.parametrize("m,want_fused", [(4096, True), (4095, False)])
def test_default_takes_the_one_pass_from_2_27_entries(m, want_fused):
    """The default mode runs the one-pass kernel once M holds 2^27 entries
    (here 4096 x 32,768) and two passes below; entries in -1..1, so every FP32
    sum is an integer below 2^24 and Q must equal the FP64 product exactly."""
    n = 32768
    g = torch.Generator(DEV).manual_seed(m)
    Md = torch.randint(-1, 2, (m, n), device=DEV, generator=g).float()
    X = torch.randint(-1, 2, (n, 8), device=DEV, generator=g).float()
    M64 = Md.double()
    ref = M64.t() @ (M64 @ X.double())
    del M64
    ctx = _lib.Context(0)
    Q, launches = _apply(ctx, Md, X, DEFAULT)
    print(f"default mode m={m} n={n}: one-pass launches {launches}", flush=True)
    assert (launches > 0) == want_fused
    assert torch.equal(Q.double(), ref), (Q.double() - ref).abs().max().item()
    del Md, ref, Q
    torch.cuda.empty_cache()


def test_one_pass_calls_on_two_streams_at_once():
    """Two contexts on two streams, each called from a host thread of its own,
    each forcing the one-pass kernel on a grid of one block per CU (n = 16,384:
    16 slices x 16 row groups): the library orders such launches one after
    another across streams (the blocks of one grid wait for each other, so two
    grids must not split the CUs); every result is exact (entries -1..1)."""
    import threading
    m, n, reps = 2048, 16384, 6
    g = torch.Generator(DEV).manual_seed(21)
    Md = torch.randint(-1, 2, (m, n), device=DEV, generator=g).float()
    Xd = [torch.randint(-1, 2, (n, 8), device=DEV, generator=g).float() for _ in range(2)]
    M64 = Md.double()
    refs = [M64.t() @ (M64 @ X.double()) for X in Xd]
    del M64
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(DEV) for _ in range(2)]
    ctxs = [_lib.Context(0) for _ in range(2)]
    outs, launches, errs = [[], []], [0, 0], []

    def run(i):
        try:
            with torch.cuda.stream(streams[i]):
                ctxs[i].set_param(_lib.MFS_PARAM_FUSED, FORCE)
                ctxs[i].track([_lib.MFS_K_MATVEC_FUSED])
                for _ in range(reps):
                    outs[i].append(ctxs[i].dense_normal_apply(Md, Xd[i]))
                launches[i] = ctxs[i].tracked(_lib.MFS_K_MATVEC_FUSED)[0]
                streams[i].synchronize()
        except Exception as e:  # noqa: BLE001 - reported below
            errs.append(e)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    print(f"two streams: one-pass launches {launches}, errors {errs}", flush=True)
    assert not errs, errs
    assert launches == [reps, reps]
    for i in range(2):
        for Q in outs[i]:
            assert torch.equal(Q, refs[i]), (i, (Q - refs[i]).abs().max().item())


def test_normal_apply_matches_fp64_ragged_rows():
    sc = scenes.coins(seed=0, n_bodies=8, pts_per_body=512, k=8)
    _, Md, _ = _setup(sc, extra_rows=45)                              # m = 4141: last panel of 13 rows
    X = torch.randn((sc.n, 8), device=DEV, generator=torch.Generator(DEV).manual_seed(0))
    _check_apply(Md, X)


def test_normal_apply_partial_slice_padded_batch():
    sc = scenes.ellipsoids(seed=1, n_bodies=4, pts_per_body=500, k=5, box=8.0)   # n = 2000
    _, Md, _ = _setup(sc)
    X = torch.randn((sc.n, 5), device=DEV, generator=torch.Generator(DEV).manual_seed(1))
    _check_apply(Md, X)


def test_normal_apply_other_widths_use_two_passes():
    sc = scenes.coins(seed=3, n_bodies=3, pts_per_body=161, k=8, box=6.0)    # n = 483
    _, Md, _ = _setup(sc)
    X = torch.randn((sc.n, 8), device=DEV, generator=torch.Generator(DEV).manual_seed(2))
    _check_apply(Md, X, want_fused=False)
    sc = scenes.coins(seed=3, n_bodies=4, pts_per_body=256, k=4)      # n = 1024: 2 slices
    _, Md, _ = _setup(sc)
    X = torch.randn((sc.n, 4), device=DEV, generator=torch.Generator(DEV).manual_seed(3))
    _check_apply(Md, X, want_fused=False)                             # batch of 4: two passes
    X = torch.randn((sc.n, 8), device=DEV, generator=torch.Generator(DEV).manual_seed(5))
    _check_apply(Md, X, want_fused=True)                              # batch of 8: one pass


def test_fused_matches_two_pass_coins_ragged_rows():
    sc = scenes.coins(seed=0, n_bodies=8, pts_per_body=512, k=8)    # n = 4096: 4 slices
    ctx, Md, B = _setup(sc, extra_rows=45)                           # m = 4141
    _compare(ctx, Md, B)


def test_fused_matches_two_pass_partial_slice_and_padded_batch():
    sc = scenes.ellipsoids(seed=1, n_bodies=4, pts_per_body=500, k=5, box=8.0)  # n = 2000
    ctx, Md, B = _setup(sc)
    _compare(ctx, Md, B)


def test_fused_source_space_right_hand_sides():
    sc = scenes.coins(seed=2, n_bodies=8, pts_per_body=256, k=1)     # n = 2048: 2 slices
    ctx, Md, _ = _setup(sc)
    rng = np.random.default_rng(7)
    G = torch.from_numpy(rng.standard_normal((sc.n, 8)).astype(np.float32)).to(DEV)
    _compare(ctx, Md, G, rhs_space=_lib.MFS_RHS_SOURCE)
    # the FP64 residual ||b - M^T M x|| / ||b|| of the solver's FP64 iterate
    # (the FP32 output adds its own rounding, amplified by the conditioning)
    s = ctx.stats()
    assert all(s.relres[c] <= 1e-6 for c in range(8))


def test_fused_not_used_for_an_odd_column_count():
    sc = scenes.coins(seed=3, n_bodies=3, pts_per_body=161, k=8, box=6.0)  # n = 483
    ctx, Md, B = _setup(sc)
    _compare(ctx, Md, B, want_fused=False)


def test_fused_full_workload_one_slice_per_cu():
    """64 ellipsoids x 1024 points, 8 right-hand sides: M is 65,536 x 65,536
    (17.2 GB), 64 slices x 4 row groups: the default takes the one pass here."""
    sc = scenes.ellipsoids(seed=0, n_bodies=64, k=8)
    ctx = _lib.Context(0)
    x = torch.from_numpy(sc.x).to(DEV)
    nx = torch.from_numpy(sc.nx).to(DEV)
    y = torch.from_numpy(sc.y).to(DEV)
    Md = ctx.assemble_dense(x, nx, y)
    B = torch.from_numpy(sc.b.astype(np.float32)).to(DEV)
    ctx.plan(sc.y, 3.0)
    ctx.factor_dense(Md, _lib.MFS_FP32)
    Xa, la, ita = _solve(ctx, Md, B, DEFAULT)
    X2, l2, it2 = _solve(ctx, Md, B, OFF)
    print(f"fused full: launches {la}, iters {ita} vs {it2}", flush=True)
    assert la > 0 and l2 == 0
    assert ctx.normal_residual(Md, B, Xa).max() <= 1e-6
    assert max(ita) <= max(it2) + 1, (ita, it2)
    X = torch.randn((sc.n, 8), device=DEV, generator=torch.Generator(DEV).manual_seed(4))
    ref = _ref_normal(Md, X)
    for mode in (DEFAULT, FORCE, OFF):
        Q, launches = _apply(ctx, Md, X, mode)
        assert (launches > 0) == (mode != OFF)
        err = ((Q - ref).norm(dim=0) / ref.norm(dim=0)).max().item()
        print(f"normal apply full mode {mode}: max rel. error {err:.2e}", flush=True)
        assert err < 2e-6, (mode, err)
    del Md, ref
    torch.cuda.empty_cache()
