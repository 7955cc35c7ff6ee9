"""CPU oracle of the caller steps either side of the solve (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker; the product path is libmfs_core.so.

FP64 torch-on-CPU restatement of fastMBCUI's geometry functions, written from
their definitions, with autograd supplying the gradients exactly as the
reference obtains them (L.backward(), core/KTAB_Mfs_spclf1d.py:136-139):

  solid_angle        utils/torch_utils.py:306-334  (compute_solid_angle)
  single_layer       utils/torch_utils.py:254-276  (compute_single_layer)
  face_centers       utils/torch_utils.py:49-60    (compute_face_centers)
  fl_matrix          utils/torch_utils.py:62-75    (compute_FL_matrix)
  quadrature         utils/torch_utils.py:336-347  (compute_quadrature)
  vertex_normals     utils/torch_utils.py:4-47     (compute_normal_vertex)
  jacobian_loop      core/KTAB_Mfs_spclf1d.py:155-163 (one backward() per entry)
  kirchhoff          core/KTAB_Mfs_spclf1d.py:97-128, the strengths from a dense
                     least-squares solve of M X = F (its renew_grad=False branch,
                     :119-121, solves with lstsq as well)

Pinned in tests/test_caller_oracle.py by closed forms, not by retyping:
the solid angle of a flat triangle over an octant (pi/2), of a square on its
axis (4 asin(a^2 / (a^2 + 4 d^2))), Gauss's theorem on closed meshes (4 pi
inside, 0 outside), the single layer of one source (1/r), and every gradient
against central finite differences.  The reference ships no vectors for these
functions.
"""
from __future__ import annotations

import numpy as np
import torch

_F64 = torch.float64


def solid_angle(V, F, S):
    """Omega (nf x n), the reference's expression (torch_utils.py:306-334) in FP64."""
    F = torch.as_tensor(np.asarray(F) if not isinstance(F, torch.Tensor) else F).long()
    R1, R2, R3 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    s = S.unsqueeze(1)
    a, b, c = R1 - s, R2 - s, R3 - s
    num = (torch.cross(b, c, dim=2) * a).sum(2)
    l1, l2, l3 = a.norm(dim=2), b.norm(dim=2), c.norm(dim=2)
    den = l1 * l2 * l3 + (a * b).sum(2) * l3 + (a * c).sum(2) * l2 + (b * c).sum(2) * l1
    return (2 * torch.atan2(num, den)).transpose(0, 1)


def single_layer(S, C):
    """SL (nf x n), SL[f][s] = 1 / |C_f - S_s| (torch_utils.py:254-276)."""
    return 1.0 / (C.unsqueeze(1) - S.unsqueeze(0)).norm(dim=2)


def face_centers(V, F):
    F = F.long()
    return (V[F[:, 0]] + V[F[:, 1]] + V[F[:, 2]]) / 3.0


def fl_matrix(V, F, Cr):
    F = F.long()
    v0, v1, v2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    normal = torch.cross(v1 - v0, v2 - v0, dim=1) / 2.0
    return torch.cat((torch.cross(Cr, normal, dim=1), normal), dim=1)


def quadrature(F, V, C):
    F = F.long()
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    s = torch.cross(p1 - p0, p2 - p0, dim=1) / 2
    return torch.cat((torch.cross(C, s, dim=1), s), dim=1).T


def vertex_normals(V, F):
    """Angle-weighted vertex normals (torch_utils.py:4-47), numpy FP64."""
    V = np.asarray(V, np.float64)
    F = np.asarray(F, np.int64)
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e1, e2, e3 = p1 - p0, p2 - p0, p2 - p1
    fn = np.cross(e1, e2)
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)

    def angle(a, b):
        d = np.einsum("ij,ij->i", a, b) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))
        return np.arccos(np.clip(d, -1.0, 1.0))

    a0 = angle(e1, e2)
    a1 = angle(-e1, e3)
    a2 = np.pi - a0 - a1
    acc = np.zeros_like(V)
    for i, a in enumerate((a0, a1, a2)):
        np.add.at(acc, F[:, i], fn * a[:, None])
    return acc / np.linalg.norm(acc, axis=1, keepdims=True)


def iq_matrix(num_links):
    """Block diagonal of [[0, I3], [I3, 0]] (KTAB:110-112)."""
    iq6 = np.block([[np.zeros((3, 3)), np.eye(3)], [np.eye(3), np.zeros((3, 3))]])
    return torch.as_tensor(np.kron(np.eye(num_links), iq6), dtype=_F64)


def kirchhoff(V_s, F_s, S_s, y_s, Qp, rho_fluid):
    """Mf and Ks_ of KTAB:97-128 for bodies already placed in the world frame,
    strengths from the dense least-squares solution of M X = F."""
    nl = len(V_s)
    S = torch.cat(tuple(S_s), 0)
    C_s = [face_centers(V, F) for V, F in zip(V_s, F_s)]
    Cr_s = [C - y for C, y in zip(C_s, y_s)]
    FL_s = [fl_matrix(V, F, Cr) for V, F, Cr in zip(V_s, F_s, Cr_s)]
    rows = []
    for i, FLi in enumerate(FL_s):
        blocks = [FLi if j == i else torch.zeros((FLi.shape[0], 6), dtype=_F64) for j in range(nl)]
        rows.append(torch.cat(blocks, 1))
    FL = torch.cat(rows, 0)
    IQ = iq_matrix(nl)
    Fm = FL @ IQ @ Qp
    M = torch.cat([solid_angle(V, F, S) for V, F in zip(V_s, F_s)], 0)
    Qm, Rm = torch.linalg.qr(M)          # least squares by QR: stable and differentiable
    strengths = torch.linalg.solve_triangular(Rm, Qm.T @ Fm, upper=True)
    phi_s = [single_layer(S, C) @ strengths for C in C_s]
    Q_s = [quadrature(F, V, Cr) for F, V, Cr in zip(F_s, V_s, Cr_s)]
    Ks = torch.cat([Q @ phi * rho_fluid for Q, phi in zip(Q_s, phi_s)], 0)
    Mf = Qp.T @ IQ @ Ks
    return Mf, Ks, strengths


def jacobian_loop(out, q):
    """d out / d q entry by entry, as the caller does: out[i, j].backward()
    then q.grad (KTAB:155-163).  Returns out.shape + q.shape."""
    J = torch.zeros(tuple(out.shape) + tuple(q.shape), dtype=q.dtype, device=q.device)
    for idx in np.ndindex(*out.shape):
        if q.grad is not None:
            q.grad.zero_()
        out[idx].backward(retain_graph=True)
        if q.grad is not None:
            J[idx] = q.grad.clone()
    if q.grad is not None:
        q.grad.zero_()
    return J
