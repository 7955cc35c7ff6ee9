"""The band form of the dense FP32 factor (k_band_gram) through the C-ABI.

The Grams of supernodes whose columns lie inside a band of 1024 columns come
straight from row-major M (32-row chunks staged in LDS, BF16 split products,
partial Grams per row part summed in part order); the other supernodes read a
compact M^T of their columns.  Against the oracle's FP64 factor (a column of
L within 5e-5 relative, as for the Gram from M^T) and against the same factor
with the band form switched off, on:

* several bodies of 512 points (m = 3072: several row parts);
* ragged row counts (m not a multiple of the 32-row chunk), with rows
  appended to M;
* a tile set without band tile counts (every supernode takes the rest path)
  and close bodies (supernodes across bands, compact M^T);
* the solve on top of it reaching its tolerance.
"""
import numpy as np
import pytest
import torch

import oracle
from fastmbcui_amd import _lib, scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)
ON, OFF = 0, 1


def _col_rel_err(pl, a, b):
    worst = 0.0
    for j in range(pl.n):
        s, e = pl.sptr[j], pl.sptr[j + 1]
        worst = max(worst, np.linalg.norm(a[s:e] - b[s:e]) / np.linalg.norm(b[s:e]))
    return worst


def _factor(M, y, bands, tile_set=0):
    ctx = _lib.Context(0)
    ctx.set_param(_lib.MFS_PARAM_FACTOR_BANDS, bands)
    ctx.set_param(_lib.MFS_PARAM_TILE_SET, tile_set)
    ctx.plan(y, 3.0)
    Md = torch.from_numpy(M).to(DEV)
    ctx.factor_dense(Md, _lib.MFS_FP32)
    s = ctx.stats()
    return ctx, Md, ctx.factor_export(), s


def _check(M, y, tile_set=0, want_band=True, solve_b=None, tol=5e-5):
    pl = oracle.plan(y, 3.0)
    ref = oracle.factor(oracle.permuted_colmajor(M.astype(np.float64), pl.Pinv), M.shape[0], pl)
    ctx, Md, val, st = _factor(M, y, ON, tile_set)
    _, _, val_off, st_off = _factor(M, y, OFF, tile_set)
    err = _col_rel_err(pl, val, ref)
    err_off = _col_rel_err(pl, val_off, ref)
    print(f"m={M.shape[0]} n={M.shape[1]} tiles {tile_set:#x}: band supernodes {st.band_supernodes} of "
          f"{st.n_super}, retried {st.retried_supernodes}; err band {err:.2e}, without {err_off:.2e}; "
          f"factor {st.factor_ms:.2f} ms (without {st_off.factor_ms:.2f} ms)", flush=True)
    assert (st.band_supernodes > 0) == want_band
    assert st_off.band_supernodes == 0
    assert err < tol and err_off < tol
    assert err <= 1.5 * err_off + 1e-6, (err, err_off)  # no worse than the Gram from M^T
    if solve_b is not None:
        B = torch.from_numpy(solve_b.astype(np.float32)).to(DEV)
        X = ctx.solve_mixed(Md, B, maxit=500, tol=1e-6)
        assert ctx.normal_residual(Md, B, X).max() <= 1e-6
    return st


def test_bands_several_bodies_and_row_parts():
    sc = scenes.ellipsoids(seed=1, n_bodies=6, pts_per_body=512, k=3, box=10.0)   # m = n = 3072
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y).astype(np.float32)
    _check(M, sc.y, solve_b=sc.b)


@pytest.mark.parametrize("extra", [0, 17, 50])
def test_bands_ragged_rows(extra):
    """n = 48 sources; m = 48, 65, 98 rows (two to four 32-row chunks, the
    last partial); appended rows are half-scaled copies of existing ones.  The
    FP32 bound of such small, poorly conditioned Grams is the one of the
    factor's other small-m test."""
    sc = scenes.coins(seed=4, n_bodies=2, pts_per_body=24, k=1, box=4.0)
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y).astype(np.float32)
    if extra:
        rng = np.random.default_rng(extra)
        M = np.vstack([M, 0.5 * M[rng.choice(M.shape[0], extra, replace=extra > M.shape[0])]]).astype(np.float32)
    _check(M, sc.y, tol=5e-2)


def test_bands_close_bodies_and_rest_path():
    """Eight coins close together: coarse supernodes take columns of several
    bodies (the rest path over a compact M^T), and bands open at columns that
    are odd before rounding down (the column pairs the band Gram loads must
    start 8-byte aligned: a misaligned pair read wrong values, 0.68 error
    against the oracle, when they were not rounded); a tile set of 16 tiles only
    leaves no band tile count, so every supernode takes the rest path.  The
    FP32 factor of this scene is 1.2e-2 from the FP64 one with or without
    bands (the Grams' conditioning), so the bound is the small-m one."""
    sc = scenes.coins(seed=0, n_bodies=8, pts_per_body=256, k=2, box=6.0)
    M = scenes.kernel_matrix(sc.x, sc.nx, sc.y).astype(np.float32)
    st = _check(M, sc.y, solve_b=sc.b, tol=5e-2)
    assert st.band_supernodes < st.n_super  # some supernodes span two bands
    _check(M, sc.y, tile_set=0x01, want_band=False, tol=5e-2)


def test_band_parameter_is_checked():
    ctx = _lib.Context(0)
    with pytest.raises(_lib.MfsError):
        ctx.set_param(_lib.MFS_PARAM_FACTOR_BANDS, 2)
