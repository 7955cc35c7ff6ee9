"""Learned warm start: a per-point MLP trained on the method's own solves.

The reference starts every PCG from zero (src/kl_chol.cu:938-941).  Here a
small network (include/mfs_mlp.h, kernels in csrc/k_mlp.hip) proposes the
initial strengths of a solve from the geometry and the bodies' velocities;
the solve then runs to the same residual tolerance as from any other start,
so the network can only change the iteration count, never the accepted answer.

* ``Network`` -- the C-ABI network: features, training steps, inference.
* ``generate(...)`` -- (geometry, v, w, strengths) samples along rigid-motion
  trajectories of coin scenes, solved by the points form.
* ``train(...)`` -- Adam over shuffled samples.
* ``guess(net, rs)`` -- initial strengths for the current pose of a RigidScene.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .scenes import RigidScene, mixed_bodies


class Network:
    def __init__(self, device=0, seed=0):
        import torch
        self.device = torch.device("cuda", device)
        h = ctypes.c_void_p()
        code = _lib.lib().mfs_mlp_create(ctypes.byref(h), self.device.index,
                                         _lib._stream_handle(self.device), int(seed))
        if code != _lib.MFS_OK:
            raise _lib.MfsError(code, "mfs_mlp_create")
        self._h = h
        self.n_params = int(_lib.lib().mfs_mlp_param_count(h))
        self.n_feat = int(_lib.lib().mfs_mlp_feature_dim())
        # strengths are learned in units of `scale` (the training targets'
        # RMS): the network's outputs start at O(1), the strengths need not
        self.scale = 1.0

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().mfs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code, what):
        if code != _lib.MFS_OK:
            raise _lib.MfsError(code, f"{what}: {_lib.lib().mfs_mlp_last_error(self._h).decode()}")

    def _stream(self):
        _lib.lib().mfs_mlp_set_stream(self._h, _lib._stream_handle(self.device))

    def params(self):
        out = np.empty(self.n_params, np.float32)
        self._stream()
        self._check(_lib.lib().mfs_mlp_get_params(self._h, out.ctypes.data), "get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self._stream()
        self._check(_lib.lib().mfs_mlp_set_params(self._h, p.ctypes.data), "set_params")

    def grads(self):
        out = np.empty(self.n_params, np.float32)
        self._stream()
        self._check(_lib.lib().mfs_mlp_get_grads(self._h, out.ctypes.data), "get_grads")
        return out

    def features(self, rs: RigidScene, b=None):
        """(features n x 20, q n x 6) for the current pose of rs (device FP32)."""
        import torch
        sc = rs.scene()
        dev = self.device
        bval = sc.b[:, 0] if b is None else b
        body = rs.body_of_point.astype(np.int32)
        starts = np.searchsorted(body, np.arange(rs.centres.shape[0] + 1)).astype(np.int64)
        vw = np.concatenate([rs.v, rs.w], axis=1)
        t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in
             (sc.y, sc.nx, bval.astype(np.float64), body, starts, rs.rot.reshape(-1, 9), rs.centres, vw)]
        n = sc.n
        feat = torch.empty((n, self.n_feat), dtype=torch.float32, device=dev)
        q = torch.empty((n, 6), dtype=torch.float32, device=dev)
        self._stream()
        self._check(_lib.lib().mfs_mlp_features(self._h, *[_lib._ptr(a) for a in t], n, _lib._ptr(feat),
                                                _lib._ptr(q)), "features")
        return feat, q

    def loss_grad(self, feat, q, target, weight):
        loss = ctypes.c_double(0)
        self._stream()
        self._check(_lib.lib().mfs_mlp_loss_grad(self._h, _lib._ptr(feat), _lib._ptr(q), _lib._ptr(target),
                                                 feat.shape[0], float(weight), ctypes.byref(loss)), "loss_grad")
        return loss.value

    def train_step(self, feat, q, target, weight, lr):
        loss = ctypes.c_double(0)
        self._stream()
        self._check(_lib.lib().mfs_mlp_train_step(self._h, _lib._ptr(feat), _lib._ptr(q), _lib._ptr(target),
                                                  feat.shape[0], float(weight), float(lr), ctypes.byref(loss)),
                    "train_step")
        return loss.value

    def predict(self, feat, q, fused=True):
        import torch
        x = torch.empty(feat.shape[0], dtype=torch.float32, device=feat.device)
        fn = _lib.lib().mfs_mlp_infer if fused else _lib.lib().mfs_mlp_predict
        self._stream()
        self._check(fn(self._h, _lib._ptr(feat), _lib._ptr(q), feat.shape[0], _lib._ptr(x)), "predict")
        return x


@dataclass
class Samples:
    feat: object      # (S * n) x 20, device
    q: object         # (S * n) x 6
    x: object         # (S * n) strengths
    n: int            # points per sample
    count: int
    solve_ms: float
    t32: object = None  # S x n x 12 packed targets (boundary-fit loss)
    s32: object = None  # S x n x 8 packed sources
    b: object = None    # S x n boundary data (FP32)
    mtb: object = None  # S x n  M^T b (FP32)


def _pack(x, nx, y):
    """FP32 point records of the matrix-free kernels (hi/lo coordinates)."""
    import torch
    xh = x.float()
    yh = y.float()
    z = torch.zeros_like(xh)
    t32 = torch.cat([xh, (x - xh.double()).float(), nx.float(), z], dim=1).contiguous()
    s32 = torch.cat([yh, (y - yh.double()).float(), z[:, :2]], dim=1).contiguous()
    return t32, s32


def coin_trajectory(seed, n_bodies=8, pts_per_body=512, box=10.0):
    """Config-1-style scene in rigid motion: 8 coins x 512 points, v, w ~ N(0, 1)."""
    return mixed_bodies(seed=seed, n_bodies=n_bodies, pts_per_body=pts_per_body, box=box, speed=1.0,
                        kinds=(0,))


def generate(net: Network, seeds, steps, every=1, dt=1e-2, rho=3.0, tol=1e-6, n_bodies=8,
             pts_per_body=512, box=10.0):
    """Solve every `every`-th pose of `steps`-step trajectories (points form,
    warm-started from the previous solve) and keep (features, q, strengths)."""
    import torch
    feats, qs, xs, t32s, s32s, bs, mtbs = [], [], [], [], [], [], []
    t_solve = 0.0
    ctx = _lib.Context(net.device.index)
    dev = net.device
    n = n_bodies * pts_per_body
    for seed in seeds:
        rs = coin_trajectory(seed, n_bodies, pts_per_body, box)
        X = None
        for step in range(steps):
            if step:
                rs.advance(dt)
            if step % every:
                continue
            sc = rs.scene()
            t0 = time.perf_counter()
            x, nx, y = (torch.from_numpy(a).to(dev) for a in (sc.x, sc.nx, sc.y))
            B = torch.from_numpy(sc.b.astype(np.float32)).to(dev)
            ctx.plan(sc.y, rho)
            ctx.factor_points(x, nx, y, _lib.MFS_FP32)
            X = ctx.solve_points(B, tol=tol, x0=X)
            t_solve += time.perf_counter() - t0
            f, q = net.features(rs)
            feats.append(f)
            qs.append(q)
            xs.append(X[:, 0].clone())
            t32, s32 = _pack(x, nx, y)
            t32s.append(t32)
            s32s.append(s32)
            bs.append(B[:, 0].clone())
            mtbs.append(ctx.points_apply(B, trans=True)[:, 0].clone())
    return Samples(torch.cat(feats), torch.cat(qs), torch.cat(xs), n, len(xs), t_solve * 1e3,
                   torch.stack(t32s), torch.stack(s32s), torch.stack(bs), torch.stack(mtbs))


def train(net: Network, data: Samples, epochs=50, batch_samples=8, lr=1e-3, seed=0, log=None):
    """Adam over shuffled minibatches of whole samples; loss = squared error
    relative to the minibatch's squared strengths.  Targets are taken in units
    of their RMS (kept as net.scale for the guesses)."""
    import torch
    rng = np.random.default_rng(seed)
    n = data.n
    net.scale = float(torch.sqrt((data.x.double() ** 2).mean()).item()) or 1.0
    losses = []
    for ep in range(epochs):
        order = rng.permutation(data.count)
        tot = 0.0
        for b0 in range(0, data.count, batch_samples):
            idx = order[b0:b0 + batch_samples]
            rows = (torch.from_numpy(idx).to(data.feat.device)[:, None] * n +
                    torch.arange(n, device=data.feat.device)[None, :]).reshape(-1)
            f = data.feat.index_select(0, rows).contiguous()
            q = data.q.index_select(0, rows).contiguous()
            t = (data.x.index_select(0, rows) / net.scale).contiguous()
            w = 1.0 / max(float((t.double() ** 2).sum()), 1e-30)
            tot += net.train_step(f, q, t, w, lr) * len(idx)
        losses.append(tot / data.count)
        if log:
            log(f"epoch {ep + 1}/{epochs}: relative loss {losses[-1]:.4e}")
    return losses


def train_fit(net: Network, data: Samples, epochs=50, batch_samples=8, lr=3e-4, seed=0, log=None,
              normal=True, lr_final_frac=0.1):
    """Adam on the data-space loss of the guesses over minibatches of whole
    samples, M applied matrix-free on the GPU: the normal-equation residual
    sum ||M^T (M x_pred - b)||^2 / sum ||M^T b||^2 (normal=True), which is
    where PCG starts from, or the boundary misfit sum ||M x_pred - b||^2 /
    sum ||b||^2."""
    import torch
    rng = np.random.default_rng(seed)
    n = data.n
    if net.scale == 1.0:
        net.scale = float(torch.sqrt((data.x.double() ** 2).mean()).item()) or 1.0
    losses = []
    elt32 = data.t32.element_size() * n * data.t32.shape[2]
    els32 = data.s32.element_size() * n * data.s32.shape[2]
    elb = data.b.element_size() * n
    for ep in range(epochs):
        order = rng.permutation(data.count)
        tot = 0.0
        lr_ep = lr * lr_final_frac ** (ep / max(epochs - 1, 1))   # exponential decay
        for b0 in range(0, data.count, batch_samples):
            idx = order[b0:b0 + batch_samples]
            S = len(idx)
            rows = (torch.from_numpy(idx).to(data.feat.device)[:, None] * n +
                    torch.arange(n, device=data.feat.device)[None, :]).reshape(-1)
            f = data.feat.index_select(0, rows).contiguous()
            q = data.q.index_select(0, rows).contiguous()
            offs = np.arange(S + 1, dtype=np.int64) * n
            ms = np.full(S, n, dtype=np.int64)
            T = (ctypes.c_void_p * S)(*[data.t32.data_ptr() + int(i) * elt32 for i in idx])
            Sr = (ctypes.c_void_p * S)(*[data.s32.data_ptr() + int(i) * els32 for i in idx])
            Bp = (ctypes.c_void_p * S)(*[data.b.data_ptr() + int(i) * elb for i in idx])
            sel = torch.from_numpy(idx).to(data.b.device)
            ref = data.mtb if normal else data.b
            w = 1.0 / max(float((ref[sel].double() ** 2).sum()), 1e-30)
            loss = ctypes.c_double(0)
            net._stream()
            net._check(_lib.lib().mfs_mlp_fit_step(net._h, _lib._ptr(f), _lib._ptr(q), S, offs.ctypes.data, T,
                                                   Sr, ms.ctypes.data, Bp, float(net.scale), float(w),
                                                   float(lr_ep), int(bool(normal)), ctypes.byref(loss)),
                       "fit_step")
            tot += loss.value * S
        losses.append(tot / data.count)
        if log:
            log(f"epoch {ep + 1}/{epochs}: relative boundary misfit {losses[-1]:.4e}")
    return losses


def guess(net: Network, rs: RigidScene):
    """Initial strengths (n x 1, FP32 device) for the current pose of rs."""
    feat, q = net.features(rs)
    return (net.predict(feat, q) * net.scale)[:, None]
