"""Pins the CPU oracle (oracle/) against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors, and none of its code
builds here (its vendored Eigen lacks the Core module), so every oracle
function is checked against an independent route: brute force on small inputs,
closed forms, textbook reductions and invariants.  Each check is built so that a
plausible slip (a dropped reversal, a wrong index, a transposed operand, a
missing term) fails it; the mutant tests at the end show that for the factor.
"""
import numpy as np
import pytest

import oracle
from fastmbcui_amd import scenes


def _rand_cloud(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, (n, 3))


def _small_coins(k=2, seed=0):
    return scenes.coins(seed=seed, n_bodies=3, pts_per_body=160, k=k, box=6.0)


# ----------------------------------------------------------------------------- ordering
@pytest.mark.parametrize("n,seed", [(2, 0), (3, 1), (17, 2), (400, 3), (1500, 4)])
def test_order_is_permutation_with_ascending_scale(n, seed):
    x = _rand_cloud(n, seed)
    P, Pinv, scale, seq, dist = oracle.fast_fps(x)
    assert sorted(P.tolist()) == list(range(n))
    assert np.array_equal(P[Pinv], np.arange(n))
    assert np.all(np.diff(scale) >= 0)
    assert np.array_equal(Pinv, seq[::-1])
    assert np.array_equal(scale, dist[::-1])


def test_root_is_closest_to_bbox_corner_lowest_index_on_ties():
    # Points 3 and 5 are both at the corner distance minimum; 3 must win.
    x = np.array([[1.0, 1.0, 1.0], [0.5, 0.2, 0.9], [0.7, 0.7, 0.1], [0.0, 0.0, 0.1],
                  [0.9, 0.1, 0.3], [0.1, 0.0, 0.0], [0.3, 0.8, 0.6]])
    _, _, _, seq, _ = oracle.fast_fps(x)
    assert seq[0] == 3


@pytest.mark.parametrize("n,seed", [(300, 0), (1200, 1)])
def test_order_equals_bruteforce_maximin_on_tie_free_input(n, seed):
    x = _rand_cloud(n, seed)
    _, _, _, seq, dist = oracle.fast_fps(x)
    bseq, bdist = oracle.maximin_brute(x, seq[0])
    assert np.array_equal(seq, bseq)
    assert np.array_equal(dist, bdist)


def test_order_on_scene_equals_bruteforce():
    sc = _small_coins()
    _, _, _, seq, dist = oracle.fast_fps(sc.y)
    bseq, bdist = oracle.maximin_brute(sc.y, seq[0])
    assert np.array_equal(seq, bseq)
    assert np.array_equal(dist, bdist)


def test_order_distances_are_distances_to_earlier_pivots():
    # Holds whatever the tie order: dist[0] is the largest distance from the root,
    # dist[k] is the distance from pivot k to the earlier pivots, and pivot k is a
    # farthest remaining point (maximin).  Lattice points force many exact ties.
    x = np.round(_rand_cloud(250, 7) * 4) / 4
    x = np.unique(x, axis=0)
    _, _, _, seq, dist = oracle.fast_fps(x)
    pts = x[seq]
    d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    assert dist[0] == pytest.approx(d[0].max(), rel=1e-15)
    run = d[0].copy()
    for k in range(1, len(seq)):
        assert dist[k] == pytest.approx(run[k], rel=1e-15)
        assert dist[k] == pytest.approx(run[k:].max(), rel=1e-15)
        run = np.minimum(run, d[k])


def test_duplicate_points_are_reported():
    x = _rand_cloud(50, 3)
    x[10] = x[20]
    with pytest.raises(oracle.OracleError) as e:
        oracle.fast_fps(x)
    assert e.value.code == oracle.MO_ERR_DUPLICATE


# ----------------------------------------------------------------------------- pattern
def _pattern_numpy(xyz_new, scale, rho):
    n = xyz_new.shape[0]
    cols = []
    for j in range(n):
        d = xyz_new[j:] - xyz_new[j]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        li = rho * scale[j:]
        lj = rho * scale[j]
        lim = np.minimum(li * li, lj * lj)
        cols.append(j + np.nonzero(d2 < lim)[0])
    return cols


@pytest.mark.parametrize("rho", [1.2, 3.0, 4.0])
def test_pattern_equals_bruteforce_predicate(rho):
    sc = _small_coins()
    P, Pinv, scale, _, _ = oracle.fast_fps(sc.y)
    r32 = float(np.float32(rho))
    ptr, ind = oracle.pattern(sc.y[Pinv], scale, r32)
    cols = _pattern_numpy(sc.y[Pinv], scale, r32)
    assert ptr[-1] == sum(len(c) for c in cols)
    for j, c in enumerate(cols):
        assert np.array_equal(ind[ptr[j]:ptr[j + 1]], c)
    # lower triangular, sorted, diagonal present
    for j in range(len(cols)):
        col = ind[ptr[j]:ptr[j + 1]]
        assert col[0] == j and np.all(np.diff(col) > 0)


def test_pattern_dense_for_huge_rho():
    x = _rand_cloud(40, 5)
    _, Pinv, scale, _, _ = oracle.fast_fps(x)
    ptr, ind = oracle.pattern(x[Pinv], scale, 1e6)
    assert ptr[-1] == 40 * 41 // 2


# ----------------------------------------------------------------------------- supernodes
def _plan(sc, rho=3.0, lam=1.5, cap=2048):
    return oracle.plan(sc.y, rho, lam, cap)


@pytest.mark.parametrize("cap", [2048, 3])
def test_aggregate_invariants(cap):
    sc = _small_coins()
    pl = _plan(sc, cap=cap)
    n = pl.n
    members = [pl.sup_ind[pl.sup_ptr[s]:pl.sup_ptr[s + 1]] for s in range(pl.n_super)]
    allm = np.concatenate(members)
    assert sorted(allm.tolist()) == list(range(n))           # partition
    for s, mem in enumerate(members):
        assert np.all(pl.sup_parent[mem] == s)
        j = mem[0]
        assert j == mem.min()                                  # opener first, and smallest
        assert len(mem) <= cap
        col = pl.ind[pl.ptr[j]:pl.ptr[j + 1]]
        assert set(mem.tolist()) <= set(col.tolist())          # members are rows of the opener
        assert np.all(pl.scale[mem] <= 1.5 * pl.scale[j])      # scale rule
        # greedy maximality: an eligible row left out went to an earlier supernode,
        # unless the cap stopped the scan
        if len(mem) < cap:
            for i in col:
                if pl.scale[i] <= 1.5 * pl.scale[j] and pl.sup_parent[i] != s:
                    assert pl.sup_parent[i] < s
    # openers are visited in increasing order
    openers = pl.sup_ind[pl.sup_ptr[:-1]]
    assert np.all(np.diff(openers) > 0)


def test_super_pattern_is_union_restricted():
    sc = _small_coins()
    pl = _plan(sc)
    for s in range(pl.n_super):
        mem = pl.sup_ind[pl.sup_ptr[s]:pl.sup_ptr[s + 1]]
        union = sorted(set(np.concatenate([pl.ind[pl.ptr[j]:pl.ptr[j + 1]] for j in mem]).tolist()))
        union = np.array(union)
        for j in mem:
            col = pl.sind[pl.sptr[j]:pl.sptr[j + 1]]
            assert np.array_equal(col, union[union >= j])
            simple = pl.ind[pl.ptr[j]:pl.ptr[j + 1]]
            assert set(simple.tolist()) <= set(col.tolist())


# ----------------------------------------------------------------------------- factor
def _factor_case(seed=0):
    sc = _small_coins(seed=seed)
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y).astype(np.float32).astype(np.float64)
    pl = _plan(sc)
    Mp = oracle.permuted_colmajor(M, pl.Pinv)
    val = oracle.factor(Mp, M.shape[0], pl)
    return sc, M, pl, Mp, val


def _closed_form_violation(M, pl, val):
    """max relative deviation of each column from Theta_j^{-1} e1 / sqrt(e1^T Theta_j^{-1} e1)."""
    Mp = M[:, pl.Pinv]
    worst = 0.0
    for j in range(pl.n):
        rows = pl.sind[pl.sptr[j]:pl.sptr[j + 1]]
        A = Mp[:, rows]
        th = A.T @ A
        e = np.zeros(len(rows)); e[0] = 1.0
        u = np.linalg.solve(th, e)
        ref = u / np.sqrt(u[0])
        got = val[pl.sptr[j]:pl.sptr[j + 1]]
        worst = max(worst, np.linalg.norm(got - ref) / np.linalg.norm(ref))
    return worst


def test_factor_columns_match_closed_form():
    sc, M, pl, Mp, val = _factor_case()
    assert _closed_form_violation(M, pl, val) < 1e-8


def test_factor_diagonal_is_positive_and_first_row():
    sc, M, pl, Mp, val = _factor_case()
    assert np.all(val[pl.sptr[:-1]] > 0)


def test_dense_pattern_gives_exact_inverse():
    # Tiny scene with a huge rho: every column is full below the diagonal, so the
    # KL factor is the exact inverse Cholesky factor: L L^T = (Mp^T Mp)^{-1}.
    sc = scenes.coins(seed=1, n_bodies=2, pts_per_body=24, k=1, box=4.0)
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    pl = oracle.plan(sc.y, 1e6)
    Mp = oracle.permuted_colmajor(M, pl.Pinv)
    val = oracle.factor(Mp, M.shape[0], pl)
    n = pl.n
    L = np.zeros((n, n))
    for j in range(n):
        L[pl.sind[pl.sptr[j]:pl.sptr[j + 1]], j] = val[pl.sptr[j]:pl.sptr[j + 1]]
    Mpd = M[:, pl.Pinv]
    inv = np.linalg.inv(Mpd.T @ Mpd)
    assert np.allclose(L @ L.T, inv, rtol=1e-8, atol=1e-10 * np.abs(inv).max())
    assert np.allclose(L, np.tril(L))


def test_factor_reports_non_spd_block():
    sc = _small_coins()
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    pl = _plan(sc)
    Mp = oracle.permuted_colmajor(M, pl.Pinv)
    # A zero column makes every Gram block that holds it singular (zero pivot).
    s = int(np.argmax(pl.local_sizes()))
    first = pl.sup_ind[pl.sup_ptr[s]]
    rows = pl.sind[pl.sptr[first]:pl.sptr[first + 1]]
    Mp = Mp.reshape(pl.n, -1).copy()
    Mp[rows[1]] = 0.0
    Mp = np.ascontiguousarray(Mp.reshape(-1))
    with pytest.raises(oracle.OracleError) as e:
        oracle.factor(Mp, M.shape[0], pl)
    assert e.value.code == oracle.MO_ERR_NOT_SPD


# --- mutants: the closed-form check must reject plausible mistakes -----------
def test_closed_form_check_rejects_mutants():
    sc, M, pl, Mp, val = _factor_case()
    assert _closed_form_violation(M, pl, val) < 1e-8
    # (1) values written in descending row order (reversal dropped)
    rev = val.copy()
    for j in range(pl.n):
        a, b = pl.sptr[j], pl.sptr[j + 1]
        rev[a:b] = val[a:b][::-1]
    assert _closed_form_violation(M, pl, rev) > 1e-3
    # (2) column of Theta^{-1} without the 1/sqrt normalisation
    unnorm = val.copy()
    for j in range(pl.n):
        a = pl.sptr[j]
        unnorm[a:pl.sptr[j + 1]] *= val[a]
    assert _closed_form_violation(M, pl, unnorm) > 1e-3
    # (3) Gram of the unpermuted matrix (permutation dropped)
    pl_id = oracle.Plan(pl.n, pl.rho, np.arange(pl.n, dtype=np.int32), np.arange(pl.n, dtype=np.int32),
                        pl.scale, pl.ptr, pl.ind, pl.sup_ptr, pl.sup_ind, pl.sup_parent, pl.sptr, pl.sind)
    Mid = oracle.permuted_colmajor(M, pl_id.Pinv)
    wrong = oracle.factor(Mid, M.shape[0], pl_id)
    assert _closed_form_violation(M, pl, wrong) > 1e-3


# ----------------------------------------------------------------------------- PCG
def test_pcg_converges_to_direct_solution():
    sc, M, pl, Mp, val = _factor_case()
    r = oracle.minvf(M, sc.b, sc.y, 3.0, 500, 1e-12, pl=pl, val=val)
    xd = np.linalg.solve(M.T @ M, M.T @ sc.b)
    assert np.linalg.norm(r.x - xd) / np.linalg.norm(xd) < 1e-8
    assert oracle.normal_residual(M, sc.b, r.x) < 1e-11


def test_pcg_stops_on_reference_threshold():
    sc, M, pl, Mp, val = _factor_case()
    r = oracle.minvf(M, sc.b, sc.y, 3.0, 500, 1e-6, pl=pl, val=val)
    assert all(rr < 1e-6 for rr in r.relres)
    # one iteration fewer would not have met the threshold
    for c in range(sc.k):
        it = r.iters[c]
        rhs = sc.b[:, c]
        _, it2, rr2 = oracle.pcg(Mp, M.shape[0], pl, val, rhs, 0, it - 1, 1e-6)
        assert rr2 >= 1e-6


def test_pcg_exact_preconditioner_converges_in_one_iteration():
    sc = scenes.coins(seed=1, n_bodies=2, pts_per_body=24, k=1, box=4.0)
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y)
    pl = oracle.plan(sc.y, 1e6)
    Mp = oracle.permuted_colmajor(M, pl.Pinv)
    val = oracle.factor(Mp, M.shape[0], pl)
    _, it, rr = oracle.pcg(Mp, M.shape[0], pl, val, sc.b[:, 0], 0, 10, 1e-8)
    assert it == 1 and rr < 1e-8


def test_pcg_zero_rhs_returns_zero():
    sc, M, pl, Mp, val = _factor_case()
    x, it, rr = oracle.pcg(Mp, M.shape[0], pl, val, np.zeros(M.shape[0]), 0, 50, 1e-6)
    assert it == 0 and rr == 0.0 and np.all(x == 0)


def test_pcg_unpreconditioned_and_preconditioned_agree():
    sc, M, pl, Mp, val = _factor_case()
    b = sc.b[:, 0]
    xp, itp, _ = oracle.pcg(Mp, M.shape[0], pl, val, b, 0, 2000, 1e-11)
    xu, itu, _ = oracle.pcg(Mp, M.shape[0], pl, val, b, 0, 5000, 1e-11, preconditioned=False)
    assert itp < itu
    assert np.linalg.norm(xp - xu) / np.linalg.norm(xp) < 1e-6


def test_vpl_solves_normal_equations_in_source_space():
    sc, M, pl, Mp, val = _factor_case()
    rng = np.random.default_rng(3)
    G = rng.standard_normal((pl.n, 2))
    r = oracle.vpl(M, G, sc.y, 3.0, 500, 1e-12, pl=pl, val=val)
    xd = np.linalg.solve(M.T @ M, G)
    assert np.linalg.norm(r.x - xd) / np.linalg.norm(xd) < 1e-8


# ----------------------------------------------------------------------------- kernel
def test_kernel_is_normal_derivative_of_laplace_fundamental_solution():
    # M[i, j] = d/d(eps) G(x_i + eps n_i, y_j), G = 1 / (4 pi |x - y|), by central differences.
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (6, 3))
    nx = rng.standard_normal((6, 3))
    nx /= np.linalg.norm(nx, axis=1, keepdims=True)
    y = rng.uniform(-1, 1, (5, 3)) + 3.0
    M = scenes.kernel_matrix(x, nx, y)
    h = 1e-5
    G = lambda p: 1.0 / (4 * np.pi * np.linalg.norm(p[:, None, :] - y[None, :, :], axis=2))
    fd = (G(x + h * nx) - G(x - h * nx)) / (2 * h)
    assert np.allclose(M, fd, rtol=1e-7, atol=0)


def test_kernel_flux_through_closed_surface_is_minus_one_inside_zero_outside():
    # Gauss: the flux of grad G(., y) through a sphere is -1 for y inside, 0 outside.
    N = 20000
    k = np.arange(N)
    z = 1 - (2 * k + 1) / N
    phi = k * np.pi * (3 - np.sqrt(5))
    r = np.sqrt(1 - z * z)
    x = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    y = np.array([[0.1, -0.2, 0.15], [2.5, 0.3, -0.4]])
    flux = scenes.kernel_matrix(x, x, y).sum(axis=0) * (4 * np.pi / N)
    assert abs(flux[0] + 1) < 1e-6
    assert abs(flux[1]) < 1e-6
