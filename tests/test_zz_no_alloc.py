"""GPU check of mfs_dense_normal_apply_noalloc (static device scratch, no
allocation) against mfs_dense_normal_apply, through the C-ABI: the same
kernels with the same split counts, so the FP64 results must be identical.
Two calls in a row, the second smaller, so stale scratch contents from the
first call would show."""
import ctypes

import numpy as np
import pytest
import torch

from fastmbcui_amd import _lib

pytestmark = pytest.mark.gpu


def _case(ctx, m, n, k, seed):
    g = torch.Generator().manual_seed(seed)
    M = torch.randn(m, n, generator=g).to("cuda")
    X = torch.randn(n, k, generator=g).to("cuda")
    print(f"noalloc case m={m} n={n} k={k} seed={seed}", flush=True)
    Q_ref = ctx.dense_normal_apply(M, X)
    Q = ctx.dense_normal_apply_noalloc(M, X)
    torch.cuda.synchronize()
    assert torch.equal(Q, Q_ref), float((Q - Q_ref).abs().max())
    # and the product itself, against an FP64 evaluation
    Md = M.double()
    want = Md.T @ (Md @ X.double())
    rel = float((Q - want).norm() / want.norm())
    assert rel < 1e-5, rel


def test_noalloc_matches_dense_normal_apply_two_calls():
    ctx = _lib.Context(0)
    try:
        _case(ctx, 3001, 2048, 8, 0)   # batch of 8, float4 rows
        _case(ctx, 517, 301, 3, 1)     # smaller, padded batch, ragged columns
    finally:
        ctx.close()


def test_noalloc_single_column_and_sixteen():
    ctx = _lib.Context(0)
    try:
        _case(ctx, 1000, 640, 1, 2)
        _case(ctx, 700, 512, 13, 3)
    finally:
        ctx.close()


def test_noalloc_rejects_shapes_beyond_the_scratch():
    ctx = _lib.Context(0)
    try:
        M = torch.zeros(8, 8, device="cuda")
        X = torch.zeros(8, 8, device="cuda")
        Q = torch.zeros(8, 8, dtype=torch.float64, device="cuda")
        L = _lib.lib()
        # claimed sizes far beyond the static scratch: rejected before any launch
        rc = L.mfs_dense_normal_apply_noalloc(ctx._h, ctypes.c_void_p(M.data_ptr()), 1 << 20, 1 << 20,
                                              ctypes.c_void_p(X.data_ptr()), 8,
                                              ctypes.c_void_p(Q.data_ptr()))
        assert rc == _lib.MFS_ERR_ARG
        assert "static scratch" in L.mfs_ctx_last_error(ctx._h).decode()
        rc = L.mfs_dense_normal_apply_noalloc(ctx._h, ctypes.c_void_p(M.data_ptr()), 8, 8,
                                              ctypes.c_void_p(X.data_ptr()), 17,
                                              ctypes.c_void_p(Q.data_ptr()))
        assert rc == _lib.MFS_ERR_ARG
    finally:
        ctx.close()
