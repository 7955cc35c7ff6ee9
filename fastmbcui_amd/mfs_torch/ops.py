"""Aliases for the names of mfs_torch/ops.py:20-97 (reference)."""
from __future__ import annotations

from torch import Tensor

from ..mfs_torch_py.ops import MinvF_float_fastGrad, VpL_fast_float

__all__ = ["test_add", "Minvb", "MinvF", "VpL"]


def test_add(a: Tensor, b: Tensor) -> Tensor:
    """Performs a + b"""
    if a.shape != b.shape:
        raise RuntimeError("test_add: shapes differ")
    return a + b


def MinvF(M: Tensor, F: Tensor, Vs_s: Tensor, rho: float, pcg_max_iter: int, tol_r: float,
          timing_verbose: bool = False, solve_verbose: bool = False) -> Tensor:
    """Performs M.pinv @ F"""
    return MinvF_float_fastGrad(M, F, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose, solve_verbose)


def Minvb(M: Tensor, b: Tensor, Vs_s: Tensor, rho: float, pcg_max_iter: int, tol_r: float,
          timing_verbose: bool = False, solve_verbose: bool = False) -> Tensor:
    """Performs M.pinv @ b for a 1-D b"""
    if b.dim() != 1:
        raise RuntimeError("Minvb expects a 1-D right-hand side")
    return MinvF_float_fastGrad(M, b.unsqueeze(1), Vs_s, rho, pcg_max_iter, tol_r,
                                timing_verbose, solve_verbose).squeeze(1)


def VpL(M: Tensor, pL: Tensor, Vs_s: Tensor, rho: float, pcg_max_iter: int, tol_r: float,
        timing_verbose: bool = False, solve_verbose: bool = False) -> Tensor:
    """Performs (M^T M)^{-1} @ pL"""
    return VpL_fast_float(M, pL, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose, solve_verbose)
