"""The FP64 certified residual through the C-ABI (mfs_normal_residual).

relres_c = ||M^T (b_c - M x_c)|| / ||M^T b_c|| against numpy in FP64 on the
same inputs.  M, b and x hold small integers (|v| <= 8), so every product and
partial sum is an integer the formats hold exactly -- including M^T b, whose
FP32 products are summed in FP32 over 32-row chunks -- and the two sides can
differ only in the rounding of the final sums of squares; a wrong index or a
lost step changes the result at O(1).  The shapes reach the row pass for FP64
batches of 8 that keeps several 256-column steps of M in flight per wave
(m >= 16,384 rows): a ragged last row block (m not a multiple of 32), column
counts that end inside a 256-column step and inside the 2048-column group of
steps in flight, and a batch of 5 padded to 8; m below 16,384 takes the
load-then-compute row pass, and so do column counts that are not a
multiple of 4 and batches of 9..16 (padded to 16); tiny scenes (5 and 70
rows) check the clamped chunk edges of the M^T r pass.
"""
import numpy as np
import pytest
import torch

from fastmbcui_amd import _lib, scenes

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _relres_fp64(M, B, X):
    M64 = M.astype(np.float64)
    R = B.astype(np.float64) - M64 @ X.astype(np.float64)
    num = np.linalg.norm(M64.T @ R, axis=0)
    den = np.linalg.norm(M64.T @ B.astype(np.float64), axis=0)
    return num / den


@pytest.mark.parametrize("m,bodies,per_body,k", [
    (16411, 4, 501, 8),     # n = 2004: one group of steps, the last step ragged
    (16384, 5, 820, 5),     # n = 4100: a second group almost all past the row; 5 of 8 columns
    (9001, 4, 501, 8),      # below 16,384 rows: the load-then-compute pass
    (16411, 3, 667, 8),     # n = 2001 (not a multiple of 4): the scalar-load row pass
    (16411, 4, 501, 13),    # a batch of 13 padded to 16
    (5, 2, 4, 8),           # n = 8, 5 rows: one short row chunk
    (70, 3, 4, 3),          # n = 12, a ragged row chunk, 3 of 4 columns
])
def test_normal_residual_matches_fp64(m, bodies, per_body, k):
    sc = scenes.coins(seed=1, n_bodies=bodies, pts_per_body=per_body, k=k, box=8.0)
    n = sc.n
    rng = np.random.default_rng(m + n)
    M = rng.integers(-8, 9, (m, n)).astype(np.float32)
    B = rng.integers(-8, 9, (m, k)).astype(np.float32)
    X = rng.integers(-8, 9, (n, k)).astype(np.float32)
    ctx = _lib.Context(0)
    ctx.plan(sc.y, 3.0)
    Md = torch.from_numpy(M).to(DEV)
    ctx.factor_dense(Md, _lib.MFS_FP32)
    got = ctx.normal_residual(Md, torch.from_numpy(B).to(DEV), torch.from_numpy(X).to(DEV))
    ref = _relres_fp64(M, B, X)
    err = np.abs(got - ref) / ref
    print(f"normal residual m={m} n={n} k={k}: max rel. diff {err.max():.2e}", flush=True)
    assert err.max() < 1e-12, (got, ref)
