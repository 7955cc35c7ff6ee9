"""Link poses and motion matrix Qp from q, and the caller's Jacobians of Mf.

fastMBCUI computes, per Kirchhoff step, every link's rotation R, origin y and
6 x nv motion matrix Qp from the configuration q in torch
(core/KTAB_Mfs_spclf1d.py:192-224 with utils/kinematics_torch.py:40-163,
handle_joint), reading the joint table (type, parent, placement) from a
pinocchio model (KTAB:66-82).  Here the table is a plain list (`JointTable`),
so no pinocchio is needed, and the per-joint rules are the same:

  revolute RX/RY/RZ      R = R_r Rod(e, q),   y offset r,            S = [0; e]
  prismatic PX/PY/PZ     R = R_r,             y offset r + R_r e q,  S = [e; 0]
  cylindrical CX/CY/CZ   translation q0 then rotation q1 about e
  PlanarYZ               R = R_r Rod(e_x, q0), offset r + R (e_y q1 + e_z q2)
  PlanarYZV2             R = R_r Rod(e_x, q2), offset r + R_r (e_y q0 + e_z q1)
  FreeFlyer              offset r + R_r q[0:3], R = R_r quat(q[3:7])

and a link's Qp is its parent's transported to the link origin plus the
joint's columns S rotated into the world frame.  As in the reference, the
free flyer's rotation is rebuilt from the quaternion's values
(utils/torch_utils.py:349-369 makes a new tensor), so no gradient flows to the
quaternion.

The caller then differentiates Mf_KDetached = Qp^T R6n K_ (K_ detached) entry
by entry: nv^2 calls of backward() (KTAB:155-163).  `jacobian` gives the same
(nv, nv, nq) tensor from one batched backward pass (is_grads_batched).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

KINDS = {"RX": 0, "RY": 1, "RZ": 2, "PX": 3, "PY": 4, "PZ": 5, "CX": 6, "CY": 7, "CZ": 8,
         "PlanarYZ": 9, "PlanarYZV2": 10, "FreeFlyer": -1}
_NQ = {0: 1, 1: 1, 2: 1, 3: 1, 4: 1, 5: 1, 6: 2, 7: 2, 8: 2, 9: 3, 10: 3, -1: 7}
_NV = {0: 1, 1: 1, 2: 1, 3: 1, 4: 1, 5: 1, 6: 2, 7: 2, 8: 2, 9: 3, 10: 3, -1: 6}
_AXIS = np.eye(3)


@dataclass
class Joint:
    kind: int                       # KINDS value
    parent: int = -1                # index of the parent joint, -1 = world
    r: np.ndarray = field(default_factory=lambda: np.zeros(3))       # placement translation
    R: np.ndarray = field(default_factory=lambda: np.eye(3))         # placement rotation


class JointTable:
    """Joints in topological order (a parent precedes its children)."""

    def __init__(self, joints):
        self.joints = list(joints)
        for i, j in enumerate(self.joints):
            if j.kind not in _NQ:
                raise ValueError(f"joint {i}: unknown kind {j.kind}")
            if not (-1 <= j.parent < i):
                raise ValueError(f"joint {i}: parent {j.parent} must precede it")
        self.nq = sum(_NQ[j.kind] for j in self.joints)
        self.nv = sum(_NV[j.kind] for j in self.joints)

    def neutral(self):
        q = np.zeros(self.nq)
        i = 0
        for j in self.joints:
            if j.kind == -1:
                q[i + 6] = 1.0
            i += _NQ[j.kind]
        return q


def _rod(axis, angle, dtype, device):
    """Rotation by `angle` (tensor) about the unit `axis` (numpy)."""
    a = torch.as_tensor(axis, dtype=dtype, device=device)
    K = torch.as_tensor(np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]],
                                  [-axis[1], axis[0], 0]]), dtype=dtype, device=device)
    return (torch.cos(angle) * torch.eye(3, dtype=dtype, device=device) + torch.sin(angle) * K
            + (1 - torch.cos(angle)) * torch.outer(a, a))


def _quat_matrix(qv):
    """Rotation of the unit quaternion (x, y, z, w), rebuilt from its values."""
    x, y, z, w = (float(t) for t in qv.detach().cpu())
    R = np.array([[1 - 2 * y * y - 2 * z * z, 2 * x * y - 2 * z * w, 2 * x * z + 2 * y * w],
                  [2 * x * y + 2 * z * w, 1 - 2 * x * x - 2 * z * z, 2 * y * z - 2 * x * w],
                  [2 * x * z - 2 * y * w, 2 * y * z + 2 * x * w, 1 - 2 * x * x - 2 * y * y]])
    return torch.as_tensor(R, dtype=qv.dtype, device=qv.device)


def _joint(j, q, iq, dtype, device):
    """(offset in the parent frame, rotation in the parent frame, S (6 x nv))."""
    t = lambda a: torch.as_tensor(np.asarray(a, np.float64), dtype=dtype, device=device)
    Rr, r = t(j.R), t(j.r)
    k = j.kind
    zeros3 = np.zeros(3)
    if k in (0, 1, 2):
        e = _AXIS[k]
        return r, Rr @ _rod(e, q[iq], dtype, device), t(np.concatenate([zeros3, e]))[:, None]
    if k in (3, 4, 5):
        e = _AXIS[k - 3]
        return r + Rr @ t(e) * q[iq], Rr, t(np.concatenate([e, zeros3]))[:, None]
    if k in (6, 7, 8):
        e = _AXIS[k - 6]
        S = np.zeros((6, 2))
        S[:3, 0] = e
        S[3:, 1] = e
        return r + Rr @ t(e) * q[iq], Rr @ _rod(e, q[iq + 1], dtype, device), t(S)
    if k == 9:
        Rq = Rr @ _rod(_AXIS[0], q[iq], dtype, device)
        off = r + Rq @ (t(_AXIS[1]) * q[iq + 1] + t(_AXIS[2]) * q[iq + 2])
        lin = torch.stack([torch.stack([q[iq] * 0, -q[iq + 2], q[iq + 1]]), t(_AXIS[1]), t(_AXIS[2])], 1)
        ang = t(np.array([[1.0, 0, 0], [0, 0, 0], [0, 0, 0]]))
        return off, Rq, torch.cat([lin, ang], 0)
    if k == 10:
        Rx = _rod(_AXIS[0], q[iq + 2], dtype, device)
        off = r + Rr @ (t(_AXIS[1]) * q[iq] + t(_AXIS[2]) * q[iq + 1])
        lin = Rx.T @ t(np.array([[0, 0, 0], [1.0, 0, 0], [0, 1.0, 0]]))
        ang = t(np.array([[0, 0, 1.0], [0, 0, 0], [0, 0, 0]]))
        return off, Rr @ Rx, torch.cat([lin, ang], 0)
    # free flyer
    return r + Rr @ q[iq:iq + 3], Rr @ _quat_matrix(q[iq + 3:iq + 7]), torch.eye(6, dtype=dtype, device=device)


def poses(table: JointTable, q: torch.Tensor):
    """(R_s, y_s, Qp_s): per link the world rotation, origin and 6 x nv motion
    matrix [linear; angular] mapping generalized velocities to the link's
    world-frame twist at its origin (KTAB:192-224)."""
    dtype, device = q.dtype, q.device
    R_s, y_s, Qp_s = [], [], []
    iq = iv = 0
    for j in table.joints:
        off, Rj, S = _joint(j, q, iq, dtype, device)
        if j.parent < 0:
            yp = torch.zeros(3, dtype=dtype, device=device)
            Rp = torch.eye(3, dtype=dtype, device=device)
            Qpp = torch.zeros((6, table.nv), dtype=dtype, device=device)
        else:
            yp, Rp, Qpp = y_s[j.parent], R_s[j.parent], Qp_s[j.parent]
        y = yp + Rp @ off
        R = Rp @ Rj
        nv = S.shape[1]
        lin = Qpp[:3] + torch.linalg.cross(Qpp[3:], (y - yp)[:, None].expand(3, table.nv), dim=0)
        before = torch.zeros((6, iv), dtype=dtype, device=device)
        after = torch.zeros((6, table.nv - iv - nv), dtype=dtype, device=device)
        cols = torch.cat([before, torch.cat([R @ S[:3], R @ S[3:]], 0), after], 1)
        Qp_s.append(torch.cat([lin, Qpp[3:]], 0) + cols)
        R_s.append(R)
        y_s.append(y)
        iq += _NQ[j.kind]
        iv += nv
    return R_s, y_s, Qp_s


def r6n(R_s):
    """Block diagonal of [[R, 0], [0, R]] per link (KTAB:144-150)."""
    dtype, device = R_s[0].dtype, R_s[0].device
    blocks = []
    for R in R_s:
        Z = torch.zeros((3, 3), dtype=dtype, device=device)
        blocks.append(torch.cat([torch.cat([R, Z], 1), torch.cat([Z, R], 1)], 0))
    return torch.block_diag(*blocks)


def jacobian(out: torch.Tensor, q: torch.Tensor) -> torch.Tensor:
    """d out / d q with shape out.shape + q.shape from one batched backward
    pass; the same numbers as calling out[i, j].backward() for every entry and
    reading q.grad (KTAB:155-163).  The graph is kept (retain_graph)."""
    flat = out.reshape(-1)
    eye = torch.eye(flat.numel(), dtype=out.dtype, device=out.device)
    (g,) = torch.autograd.grad(flat, q, grad_outputs=eye, is_grads_batched=True,
                               retain_graph=True, allow_unused=True)
    if g is None:
        g = torch.zeros((flat.numel(),) + tuple(q.shape), dtype=q.dtype, device=q.device)
    return g.reshape(tuple(out.shape) + tuple(q.shape))


def mf_kdetached_jacobian(Qp: torch.Tensor, R6n: torch.Tensor, K: torch.Tensor, q: torch.Tensor):
    """(Mf_KDetached, d Mf_KDetached / d q) for Mf_KDetached = Qp^T R6n K with K
    detached (KTAB:152-163)."""
    Mf = Qp.T @ R6n @ K.detach()
    return Mf, jacobian(Mf, q)
