"""``mfs_torch`` operators: same names, schemas, checks, fake kernels and autograd
formula as fastMBCUI's (mfs_torch/csrc/mfs_torch_ops.cpp:7-10,
mfs_torch/csrc/mfs_torch_ops.cu:21-375, core/mfs_torch_py/ops.py:1-71).

Implementations are registered for the CUDA dispatch key, which is the key
HIP tensors carry under PyTorch-ROCm (the reference registers the same key,
mfs_torch_ops.cu:372-375).  They call the C-ABI of libmfs_core.so on torch's
current stream.
"""
from __future__ import annotations

import time

import torch
from torch import Tensor

from .. import _lib, runtime

__all__ = ["MinvF_float_fastGrad", "VpL_fast_float", "last_solve_info"]

_SCHEMA_ARGS = ("(Tensor M, Tensor {b}, Tensor Vs_s, float rho, int pcg_max_iter, float tol_r, "
                "bool timing_verbose=False, bool solve_verbose=False) -> Tensor")

_lib_def = torch.library.Library("mfs_torch", "DEF")
_lib_def.define("VpL_fast_float" + _SCHEMA_ARGS.format(b="pL"))
_lib_def.define("MinvF_float_fastGrad" + _SCHEMA_ARGS.format(b="F"))
_lib_impl = torch.library.Library("mfs_torch", "IMPL", "CUDA")


def last_solve_info() -> runtime.SolveInfo:
    """Iterations and residual per column of the last call (the reference drops them)."""
    return runtime.last_info


def _check_inputs(M, B, Vs_s, name_b, forward):
    # mfs_torch_ops.cu:24-33 / :202-211
    torch._check(B.dim() == 2, lambda: f"{name_b} must be a 2D tensor")
    if forward:
        torch._check(M.size(0) == B.size(0), lambda: "M and F must have the same first dimension")
    else:
        torch._check(M.size(1) == B.size(0), lambda: "M and pL must have compatible dimensions")
    torch._check(M.size(1) == Vs_s.size(0), lambda: "M and Vs_s must have compatible dimensions")
    torch._check(Vs_s.size(1) == 3, lambda: "Vs_s must be 3d points")
    torch._check(M.dtype == torch.float32, lambda: "M must be float32")
    torch._check(B.dtype == torch.float32, lambda: f"{name_b} must be float32")
    torch._check(Vs_s.dtype == torch.float32, lambda: "Vs_s must be float32")
    torch._check(M.is_cuda and B.is_cuda, lambda: "M and the right-hand side must be device tensors")
    torch._check(Vs_s.device.type == "cpu", lambda: "Vs_s must be a CPU tensor")


def _report(info, timing_verbose, solve_verbose, total_ms):
    if timing_verbose:
        # the reference's phases (mfs_torch_ops.cu:48-193); permutation, the
        # covariance copy, the cache copy and the back-permutation are fused
        # into the kernels here and take no time of their own, and a reused
        # plan or factor costs nothing
        plan = 0.0 if info.plan_reused else info.plan_ms
        order = 0.0 if info.plan_reused else info.plan_order_ms
        pattern = 0.0 if info.plan_reused else info.plan_pattern_ms
        phases = [("prep", info.prep_ms), ("fps_sampler", order), ("permutation", 0.0),
                  ("build_cov_mat", 0.0), ("pattern", pattern),
                  ("setup", max(plan - order - pattern, 0.0)),
                  ("compute", 0.0 if info.factor_reused else info.factor_ms), ("cache", 0.0),
                  ("pcg", info.solve_ms), ("perm_back", 0.0), ("total", total_ms)]
        for name, ms in phases:
            print(f"TIME({name}): {ms} ms")
    if solve_verbose:
        for c, (it, rr) in enumerate(zip(info.iters, info.relres)):
            print(f"pcg column {c}: iterations {it}, relative residual {rr:.3e}")


def _solve(M, B, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose, solve_verbose, forward):
    t0 = time.perf_counter()
    Mc = M.contiguous()
    Bc = B.contiguous()
    mode = runtime.solver_mode()
    prec = _lib.MFS_FP32 if mode == "mixed" else _lib.MFS_FP64
    prep_ms = (time.perf_counter() - t0) * 1e3
    ctx, info = runtime.prepared(Mc, Vs_s, rho, prec)
    info.prep_ms = prep_ms
    n, k = Mc.shape[1], Bc.shape[1]
    out = torch.empty((n, k), dtype=torch.float32, device=M.device)
    if k > 0:
        if mode == "mixed":
            # pcg_max_iter caps the inner iterations of all refinement sweeps
            # together, as it caps each column's PCG in the plain form
            ctx.solve_mixed(Mc, Bc, maxit=int(pcg_max_iter), tol=float(tol_r), out=out,
                            rhs_space=_lib.MFS_RHS_BOUNDARY if forward else _lib.MFS_RHS_SOURCE)
        else:
            ctx.solve_dense(Mc, Bc, _lib.MFS_RHS_BOUNDARY if forward else _lib.MFS_RHS_SOURCE,
                            int(pcg_max_iter), float(tol_r), out=out)
        s = ctx.stats()
        info.iters = [int(s.iters[c]) for c in range(min(k, 64))]
        info.relres = [float(s.relres[c]) for c in range(min(k, 64))]
        info.solve_ms = s.solve_ms
    info.mode = mode
    runtime.last_info = info
    if timing_verbose:
        torch.cuda.synchronize(M.device)
    _report(info, timing_verbose, solve_verbose, (time.perf_counter() - t0) * 1e3)
    return out


@torch.library.impl(_lib_impl, "MinvF_float_fastGrad")
def _minvf_cuda(M, F, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose=False, solve_verbose=False):
    _check_inputs(M, F, Vs_s, "F", True)
    out = _solve(M, F, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose, solve_verbose, True)
    runtime.mark_forward()
    return out


@torch.library.impl(_lib_impl, "VpL_fast_float")
def _vpl_cuda(M, pL, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose=False, solve_verbose=False):
    _check_inputs(M, pL, Vs_s, "pL", False)
    # mfs_torch_ops.cu:213-215: the backward solve needs a preceding forward.
    torch._check(runtime.forward_seen(),
                 lambda: "Float cache tensors are not defined. Please run MinvF_float first.")
    return _solve(M, pL, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose, solve_verbose, False)


# ---------------------------------------------------------------- Python surface
def MinvF_float_fastGrad(M: Tensor, F: Tensor, Vs_s: Tensor, rho: float, pcg_max_iter: int, tol_r: float,
                         timing_verbose: bool = False, solve_verbose: bool = False) -> Tensor:
    """Performs M.pinv @ F"""
    return torch.ops.mfs_torch.MinvF_float_fastGrad.default(M, F, Vs_s, rho, pcg_max_iter, tol_r,
                                                            timing_verbose, solve_verbose)


def VpL_fast_float(M: Tensor, pL: Tensor, Vs_s: Tensor, rho: float, pcg_max_iter: int, tol_r: float,
                   timing_verbose: bool = False, solve_verbose: bool = False) -> Tensor:
    """Performs (M^T M)^{-1} @ pL"""
    return torch.ops.mfs_torch.VpL_fast_float.default(M, pL, Vs_s, rho, pcg_max_iter, tol_r,
                                                      timing_verbose, solve_verbose)


@torch.library.register_fake("mfs_torch::MinvF_float_fastGrad")
def _(M, F, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose=False, solve_verbose=False):
    torch._check(M.shape[0] == F.shape[0])
    torch._check(M.shape[1] == Vs_s.shape[0])
    torch._check(Vs_s.shape[1] == 3)
    torch._check(M.dtype == torch.float)
    torch._check(F.dtype == torch.float)
    torch._check(Vs_s.dtype == torch.float)
    torch._check(M.device == F.device)
    return torch.empty([Vs_s.shape[0], F.shape[1]], dtype=torch.float, device=M.device)


@torch.library.register_fake("mfs_torch::VpL_fast_float")
def _(M, pL, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose=False, solve_verbose=False):
    torch._check(M.shape[1] == pL.shape[0])
    torch._check(M.shape[1] == Vs_s.shape[0])
    torch._check(Vs_s.shape[1] == 3)
    torch._check(M.dtype == torch.float)
    torch._check(pL.dtype == torch.float)
    torch._check(Vs_s.dtype == torch.float)
    torch._check(M.device == pL.device)
    return torch.empty([Vs_s.shape[0], pL.shape[1]], dtype=torch.float, device=M.device)


# ---------------------------------------------------------------- autograd
# A = (M^T M)^{-1} M^T F.  With V = (M^T M)^{-1} G for the incoming gradient G:
#   dL/dF = M V,   dL/dM = F V^T - M (A V^T + V A^T)
# (core/mfs_torch_py/ops.py:42-58).  The reference forms W = A V^T (n x n) and
# M (W + W^T) (m n^2 flops); here the same quantity is evaluated as
# (F - M A) V^T - (M V) A^T, M V shared with dL/dF: mfs_dense_grad reads M for
# M A and M V (once when 2k <= 8, twice above) and writes dL/dM once, as one
# rank-2k outer product.
def _setup_context(ctx, inputs, output):
    M, F, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose, solve_verbose = inputs
    if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
        saved = (M, F, Vs_s, output)
        ctx.rho = rho
        ctx.pcg_max_iter = pcg_max_iter
        ctx.tol_r = tol_r
        ctx.timing_verbose = timing_verbose
        ctx.solve_verbose = solve_verbose
    else:
        saved = (None, None, None, None)
    ctx.save_for_backward(*saved)


def _backward(ctx, grad):
    grad_M, grad_F = None, None
    if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
        M, F, Vs_s, A = ctx.saved_tensors
        VpL = torch.ops.mfs_torch.VpL_fast_float.default(M, grad.contiguous(), Vs_s, ctx.rho,
                                                         ctx.pcg_max_iter, ctx.tol_r,
                                                         ctx.timing_verbose, ctx.solve_verbose)
        if M.dtype == torch.float32 and M.is_cuda and VpL.shape[1] <= 16:
            # row passes over M for M A and M V, one pass writing grad_M
            dctx = runtime.caller_context(M.device)
            grad_M, grad_F = dctx.dense_grad(M.contiguous(), F, A, VpL, ctx.needs_input_grad[0],
                                             ctx.needs_input_grad[1])
        else:
            MV = M @ VpL
            if ctx.needs_input_grad[0]:
                grad_M = (F - M @ A) @ VpL.T - MV @ A.T
            if ctx.needs_input_grad[1]:
                grad_F = MV
    return grad_M, grad_F, None, None, None, None, None, None


torch.library.register_autograd("mfs_torch::MinvF_float_fastGrad", _backward,
                                setup_context=_setup_context)
