"""CPU oracle of the learned warm start (test infrastructure only).

numpy restatement of the network the HIP kernels implement (csrc/k_mlp.hip,
include/mfs_mlp.h): body-frame features with a 16-nearest-neighbour max-pool,
MLP 20 -> 128 -> 128 -> 128 -> 6 with ReLU, head x = out . q, loss
w * sum (x - t)^2, its gradients, and Adam.  The reference has no network
(its PCG starts from zero, src/kl_chol.cu:938-941), so this oracle is pinned
by mathematics instead of reference vectors: the analytic gradients against
central finite differences in FP64 (tests/test_mlp_oracle.py), the features
against their definition on hand-built bodies, Adam against its closed form.

Parameters are one flat FP32 vector [W0 W1 W2 W3 b0 b1 b2 b3], W_l (in x out).
The forward pass uses the weights rounded to BF16 (round to nearest even), as
the GPU does; gradients are with respect to those rounded weights.  With
bf16_ops (the default) every matrix product also takes its other operand
rounded to BF16 from FP32, as the GPU's BF16 matrix-core products do: the
layer inputs in the forward pass, and the back-propagated gradient G in both
backward products (dW = bf(A)^T bf(G), dA = bf(G) bf(W)^T); the bias
gradients and ReLU masks use the unrounded values.  The products are then
approximations of the exact gradient, which the finite-difference check pins
with bf16_ops=False.
"""
from __future__ import annotations

import numpy as np

DIMS = (20, 128, 128, 128, 6)
KNN = 16


def layout():
    w_off, b_off, off = [], [], 0
    for l in range(4):
        w_off.append(off)
        off += DIMS[l] * DIMS[l + 1]
    n_w = off
    for l in range(4):
        b_off.append(off)
        off += DIMS[l + 1]
    return w_off, b_off, n_w, off


def bf16_round(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(p, round_weights=True):
    w_off, b_off, n_w, _ = layout()
    W, b = [], []
    for l in range(4):
        w = p[w_off[l]:w_off[l] + DIMS[l] * DIMS[l + 1]].reshape(DIMS[l], DIMS[l + 1])
        W.append((bf16_round(w) if round_weights else w).astype(np.float64))
        b.append(p[b_off[l]:b_off[l] + DIMS[l + 1]].astype(np.float64))
    return W, b


def features(y, nrm, bval, body, starts, rot, ctr, vw):
    """(n x 20 features, n x 6 q) -- the definition in csrc/k_mlp.hip."""
    n = y.shape[0]
    feat = np.zeros((n, 20))
    q = np.zeros((n, 6))
    R = rot.reshape(-1, 3, 3)
    loc = lambda b, v: R[b].T @ v
    for j in range(n):
        b = body[j]
        r = loc(b, y[j] - ctr[b])
        nl = loc(b, nrm[j])
        vl = loc(b, vw[b, :3])
        wl = loc(b, vw[b, 3:])
        feat[j, 0:3], feat[j, 3:6], feat[j, 6:9], feat[j, 9:12] = r, nl, vl, wl
        feat[j, 12] = bval[j]
        q[j, :3], q[j, 3:] = vl, wl
        s0, s1 = starts[b], starts[b + 1]
        others = np.array([k for k in range(s0, s1) if k != j], dtype=np.int64)
        if others.size == 0:
            continue
        dd = np.sum((y[others] - y[j]) ** 2, axis=1).astype(np.float32)
        nb = others[np.argsort(dd, kind="stable")[:KNN]]
        rk = np.array([loc(b, y[k] - ctr[b]) for k in nb])
        nk = np.array([loc(b, nrm[k]) for k in nb])
        feat[j, 13:16] = (rk - r).max(axis=0)
        feat[j, 16:19] = nk.max(axis=0)
        feat[j, 19] = bval[nb].max()
    return feat, q


def _op(a, bf16_ops):
    """A matrix-product operand: FP32, then BF16 when bf16_ops."""
    return bf16_round(np.asarray(a, np.float32)).astype(np.float64) if bf16_ops else a


def forward(p, feat, round_weights=True, bf16_ops=True):
    """Returns (out n x 6, activations [A0 .. A3])."""
    W, b = unpack(p, round_weights)
    A = [np.asarray(feat, np.float64)]
    h = A[0]
    for l in range(4):
        z = _op(h, bf16_ops) @ W[l] + b[l]
        h = np.maximum(z, 0.0) if l < 3 else z
        if l < 3:
            A.append(h)
    return h, A


def loss_grad(p, feat, q, target, weight, round_weights=True, bf16_ops=True):
    """(loss, flat gradient) of weight * sum_j (out_j . q_j - t_j)^2."""
    W, b = unpack(p, round_weights)
    out, A = forward(p, feat, round_weights, bf16_ops)
    x = np.sum(out * q, axis=1)
    e = x - target
    loss = weight * float(np.sum(e * e))
    G = 2.0 * weight * e[:, None] * q
    w_off, b_off, _, n_all = layout()
    g = np.zeros(n_all)
    for l in range(3, -1, -1):
        Gb = _op(G, bf16_ops)
        g[w_off[l]:w_off[l] + W[l].size] = (_op(A[l], bf16_ops).T @ Gb).ravel()
        g[b_off[l]:b_off[l] + b[l].size] = G.sum(axis=0)
        if l > 0:
            G = (Gb @ W[l].T) * (A[l] > 0)
    return loss, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
