"""The caller's steps either side of the solve, on the GPU, with autograd.

fastMBCUI's Kirchhoff-tensor step (core/KTAB_Mfs_spclf1d.py:97-128) does, per
call, with every body's mesh placed in the world frame:

    S   = cat(S_s)                              sources
    M   = cat(compute_solid_angle(V, F, S))     (num_faces x n)      O(m n)
    F   = FL @ IQ @ Qp                          boundary data        O(m)
    X   = MinvF_float_fastGrad(M, F, S, ...)    the solve
    phi = compute_single_layer(S, C) @ X        per body (nf x nv)   O(m n nv)
    Ks_ = cat(Q @ phi * rho_fluid);  Mf = Qp^T @ IQ @ Ks_

and L = 0.5 v^T Mf v; L.backward() (:136-139) differentiates all of it.  The
two O(m n) steps are HIP kernels here (libmfs_core.so: mfs_solid_angle,
mfs_single_layer_apply and their backwards), wrapped as autograd functions
under the reference's names; the O(m) steps stay torch tensor expressions on
the same device.  compute_single_layer returns an operator, so the caller's
`SL @ strengths` runs the matrix-free kernel and the nf x n matrix is never
formed.  Device tensors only: a CPU tensor is an error, not a fallback.
"""
from __future__ import annotations

import numpy as np
import torch

from . import runtime

__all__ = ["compute_solid_angle", "compute_single_layer", "single_layer_apply",
           "compute_face_centers", "compute_FL_matrix", "compute_quadrature", "iq_matrix",
           "kirchhoff_tensor", "kirchhoff_step"]


def _require_device(*ts):
    for t in ts:
        if not t.is_cuda:
            raise RuntimeError("caller kernels take device tensors (HIP); got a CPU tensor")


class _SolidAngle(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, F, S):
        _require_device(V, F, S)
        c = runtime.caller_context(V.device)
        ctx.save_for_backward(V, F, S)
        return c.solid_angle(V.detach(), F, S.detach())

    @staticmethod
    def backward(ctx, gM):
        V, F, S = ctx.saved_tensors
        need_v, _, need_s = ctx.needs_input_grad
        if not (need_v or need_s):
            return None, None, None
        c = runtime.caller_context(V.device)
        gV, gS = c.solid_angle_backward(V.detach(), F, S.detach(), gM, need_v, need_s)
        if gS is not None:
            gS = gS.to(S.dtype)
        return gV, None, gS


def compute_solid_angle(V, F, S):
    """Solid angle (num_faces x num_sources) of every face of (V, F) seen from
    every source, 2 atan2(N, D) (utils/torch_utils.py:306-334).  float32 or
    float64 V (S is cast to V's dtype); F int32/int64 (num_faces x 3)."""
    return _SolidAngle.apply(V, F, S)


class _SingleLayer(torch.autograd.Function):
    @staticmethod
    def forward(ctx, S, C, X):
        _require_device(S, C, X)
        c = runtime.caller_context(X.device)
        ctx.save_for_backward(S, C, X)
        return c.single_layer_apply(S.detach(), C.detach(), X.detach())

    @staticmethod
    def backward(ctx, gphi):
        S, C, X = ctx.saved_tensors
        need_s, need_c, need_x = ctx.needs_input_grad
        if not (need_s or need_c or need_x):
            return None, None, None
        c = runtime.caller_context(X.device)
        gS, gC, gX = c.single_layer_backward(S.detach(), C.detach(), X.detach(), gphi,
                                             need_s, need_c, need_x)
        return (None if gS is None else gS.to(S.dtype), None if gC is None else gC.to(C.dtype), gX)


def single_layer_apply(S, C, X):
    """phi = SL @ X with SL[f][s] = 1 / |C[f] - S[s]| (utils/torch_utils.py:254-276
    and KTAB:123-124), without storing SL.  X: num_sources x k."""
    return _SingleLayer.apply(S, C, X)


class SingleLayer:
    """The single-layer matrix as an operator: `SL @ X` applies it matrix-free."""

    def __init__(self, S, C):
        self.S, self.C = S, C

    @property
    def shape(self):
        return (self.C.shape[0], self.S.shape[0])

    def __matmul__(self, X):
        if X.dim() == 1:
            return single_layer_apply(self.S, self.C, X.unsqueeze(1)).squeeze(1)
        return single_layer_apply(self.S, self.C, X)


def compute_single_layer(S, C):
    """compute_single_layer(S, C) (utils/torch_utils.py:254-276) as an operator:
    `compute_single_layer(S, C) @ strengths` gives what the reference's dense
    matrix product gives."""
    return SingleLayer(S, C)


def compute_face_centers(V, F):
    """(V[F0] + V[F1] + V[F2]) / 3 (utils/torch_utils.py:49-60)."""
    F = F.long()
    return (V[F[:, 0]] + V[F[:, 1]] + V[F[:, 2]]) / 3.0


def compute_FL_matrix(V, F, Cr):
    """[Cr x n, n] with n = (V1 - V0) x (V2 - V0) / 2 (utils/torch_utils.py:62-75)."""
    F = F.long()
    v0, v1, v2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    normal = torch.linalg.cross(v1 - v0, v2 - v0) / 2.0
    return torch.cat((torch.linalg.cross(Cr, normal), normal), dim=1)


def compute_quadrature(F, V, C):
    """[C x s, s]^T with s the half cross product of the edges (utils/torch_utils.py:336-347)."""
    F = F.long()
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    s = torch.linalg.cross(p1 - p0, p2 - p0) / 2
    return torch.cat((torch.linalg.cross(C, s), s), dim=1).T


def iq_matrix(num_links, dtype=torch.float32, device=None):
    """Block diagonal of [[0, I3], [I3, 0]], one block per link (KTAB:109-111)."""
    iq6 = np.block([[np.zeros((3, 3)), np.eye(3)], [np.eye(3), np.zeros((3, 3))]])
    return torch.as_tensor(np.kron(np.eye(num_links), iq6), dtype=dtype, device=device)


def kirchhoff_tensor(V_s, F_s, S_s, y_s, Qp, rho_fluid, rho_mfs, pcg_max_iter, tol_r,
                     solid_angle=None, single_layer=None):
    """Mf (nv x nv), Ks_ (6 num_links x nv) and the strengths (n x nv) of
    KTAB:97-128 for bodies already placed in the world frame (V_s, S_s: one
    (count x 3) tensor per body; y_s: body origins; Qp: 6 num_links x nv).
    Differentiable in every tensor argument; the solve is the mfs_torch op
    (MinvF_float_fastGrad), so its dtype is float32 as in the reference.
    `solid_angle(V, F, S)` and `single_layer(S, C)` replace the two O(m n)
    steps (default: the HIP kernels), e.g. to time other implementations."""
    solid_angle = solid_angle or compute_solid_angle
    single_layer = single_layer or compute_single_layer
    from .mfs_torch import ops as mfs_ops

    nl = len(V_s)
    dt, dev = V_s[0].dtype, V_s[0].device
    S = torch.cat(tuple(S_s), dim=0)
    C_s = [compute_face_centers(V, F) for V, F in zip(V_s, F_s)]
    Cr_s = [C - y for C, y in zip(C_s, y_s)]
    FL_s = [compute_FL_matrix(V, F, Cr) for V, F, Cr in zip(V_s, F_s, Cr_s)]
    FL = torch.cat([torch.cat([FLi if j == i else torch.zeros((FLi.shape[0], 6), dtype=dt, device=dev)
                               for j in range(nl)], dim=1) for i, FLi in enumerate(FL_s)], dim=0)
    IQ = iq_matrix(nl, dt, dev)
    F = FL @ IQ @ Qp
    M = torch.cat([solid_angle(V, Fi, S) for V, Fi in zip(V_s, F_s)], dim=0)
    strengths = mfs_ops.MinvF_float_fastGrad(M, F, S.detach().cpu(), rho_mfs, pcg_max_iter, tol_r)
    phi_s = [single_layer(S, C) @ strengths for C in C_s]
    Q_s = [compute_quadrature(Fi, V, Cr) for Fi, V, Cr in zip(F_s, V_s, Cr_s)]
    Ks = torch.cat([Q @ phi * rho_fluid for Q, phi in zip(Q_s, phi_s)], dim=0)
    Mf = Qp.T @ IQ @ Ks
    return Mf, Ks, strengths


def kirchhoff_step(table, bodies, q, v, rho_fluid, rho_mfs, pcg_max_iter, tol_r, grad=True):
    """One call of the reference's compute_kirchhoff_tensor (KTAB:88-169,
    without drag) on the GPU.  `table` is a kinematics.JointTable with one
    link per body; `bodies` a list of (V0, F, S0) device tensors in the link
    frames (vertices, faces, sources); q (nq) and v (nv) numpy or tensors.

    Returns a dict with Mf (nv x nv), K (6 num_links x nv, the detached K_ of
    :146-147) and, when `grad`, grad_L_q = dL/dq for L = 0.5 v^T Mf v
    (:136-139) and grad_Mf_q_KDetached = d(Qp^T R6n K_)/dq (nv x nv x nq,
    :152-163) from one batched backward pass instead of nv^2 backward() calls."""
    from . import kinematics as kin

    V0 = bodies[0][0]
    dt, dev = V0.dtype, V0.device
    q_t = torch.as_tensor(q, dtype=dt, device=dev).detach().clone().requires_grad_(grad)
    v_t = torch.as_tensor(v, dtype=dt, device=dev)
    R_s, y_s, Qp_s = kin.poses(table, q_t)
    Qp = torch.cat(tuple(Qp_s), dim=0)
    V_s = [V0_ @ R.T + y for (V0_, _, _), R, y in zip(bodies, R_s, y_s)]
    S_s = [S0 @ R.T + y for (_, _, S0), R, y in zip(bodies, R_s, y_s)]
    F_s = [F for (_, F, _) in bodies]
    Mf, Ks, _ = kirchhoff_tensor(V_s, F_s, S_s, y_s, Qp, rho_fluid, rho_mfs, pcg_max_iter, tol_r)
    IQ = iq_matrix(len(bodies), dt, dev)
    R6n = kin.r6n(R_s)
    K = (R6n.T @ IQ @ Ks).detach()
    out = dict(Mf=Mf.detach(), K=K)
    if grad:
        L = 0.5 * v_t @ Mf @ v_t
        (gq,) = torch.autograd.grad(L, q_t, retain_graph=True)
        out["grad_L_q"] = gq
        out["grad_Mf_q_KDetached"] = kin.mf_kdetached_jacobian(Qp, R6n, K, q_t)[1]
    return out
