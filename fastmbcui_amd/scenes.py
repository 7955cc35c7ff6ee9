"""Seeded synthetic MFS scenes (collocation points, normals, sources, right-hand sides).

The shipped drivers of fastMBCUI build their scenes from OBJ meshes
(core/KTAB_Mfs_spclf1d.py:33-51).  The workloads measured here are point
scenes instead:

* ``coins(seed)``: 8 cylinders (radius 1, thickness 0.1), 512 area-uniform
  surface points each (N = 4096), sources at x - 0.02 n.
* ``ellipsoids(seed)``: 64 ellipsoids with semi-axes ~ U(0.5, 1.5), 1024
  Fibonacci-mapped points each (N = 65,536), sources at x - 0.1 min(a,b,c) n.

Bodies get a uniform random rotation (normalised 4-Gaussian quaternion) and a
centre drawn uniformly in a box, rejected while its bounding sphere overlaps an
earlier body.  Right-hand sides are Neumann data of rigid motions:
``b_i = n_i . (v + w x (x_i - c_body))`` with v, w ~ N(0, I3) per body and per
right-hand side.

The kernel is the normal derivative of the Laplace fundamental solution
``G(x, y) = 1 / (4 pi |x - y|)`` at the target:
``M[i, j] = -(x_i - y_j) . n_i / (4 pi |x_i - y_j|^3)``.  The reference's
``laplace_nm_3d_sl`` (src/kernels.h:24-40) evaluates ``(x - y) . n_x / r^3``,
i.e. ``-4 pi`` times this kernel; the solve is linear, so that only rescales
the strengths.

Everything here is numpy FP64 and depends only on the seed, so a CPU check and
a GPU run see bit-identical points.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

FOUR_PI = 4.0 * np.pi


@dataclass
class Scene:
    """Targets x (m x 3), unit normals nx (m x 3), sources y (n x 3), RHS b (m x k)."""

    name: str
    x: np.ndarray
    nx: np.ndarray
    y: np.ndarray
    b: np.ndarray
    body_of_point: np.ndarray
    centres: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def m(self) -> int:
        return int(self.x.shape[0])

    @property
    def n(self) -> int:
        return int(self.y.shape[0])

    @property
    def k(self) -> int:
        return int(self.b.shape[1])


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, a, b, c = q
    return np.array([
        [1 - 2 * (b * b + c * c), 2 * (a * b - c * w), 2 * (a * c + b * w)],
        [2 * (a * b + c * w), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w)],
        [2 * (a * c - b * w), 2 * (b * c + a * w), 1 - 2 * (a * a + b * b)],
    ])


def _place(rng, radii, box, max_tries=100000):
    """Centres uniform in [0, box]^3, rejecting overlap of bounding spheres."""
    centres = []
    for r in radii:
        for _ in range(max_tries):
            c = rng.uniform(0.0, box, 3)
            if all(np.linalg.norm(c - c2) > r + r2 for c2, r2 in zip(centres, radii)):
                centres.append(c)
                break
        else:
            raise RuntimeError("could not place body without overlap")
    return np.array(centres)


def _rigid_rhs(rng, x, nx, body, centres, k):
    nb = centres.shape[0]
    b = np.empty((x.shape[0], k))
    for r in range(k):
        v = rng.standard_normal((nb, 3))
        w = rng.standard_normal((nb, 3))
        rel = x - centres[body]
        vel = v[body] + np.cross(w[body], rel)
        b[:, r] = np.einsum("ij,ij->i", nx, vel)
    return b


def _cylinder_surface(rng, npts, radius, half_t):
    """Area-uniform samples on a closed cylinder (caps + side), body frame."""
    a_side = 2 * np.pi * radius * 2 * half_t
    a_cap = np.pi * radius * radius
    probs = np.array([a_side, a_cap, a_cap]) / (a_side + 2 * a_cap)
    face = rng.choice(3, size=npts, p=probs)
    th = rng.uniform(0.0, 2 * np.pi, npts)
    u = rng.uniform(0.0, 1.0, npts)
    z = rng.uniform(-half_t, half_t, npts)
    p = np.empty((npts, 3))
    nrm = np.empty((npts, 3))
    side = face == 0
    p[side] = np.stack([radius * np.cos(th[side]), radius * np.sin(th[side]), z[side]], 1)
    nrm[side] = np.stack([np.cos(th[side]), np.sin(th[side]), np.zeros(side.sum())], 1)
    for f, sgn in ((1, 1.0), (2, -1.0)):
        sel = face == f
        rr = radius * np.sqrt(u[sel])
        p[sel] = np.stack([rr * np.cos(th[sel]), rr * np.sin(th[sel]), np.full(sel.sum(), sgn * half_t)], 1)
        nrm[sel] = np.array([0.0, 0.0, sgn])
    return p, nrm


def coins(seed: int = 0, n_bodies: int = 8, pts_per_body: int = 512, k: int = 1,
          box: float = 10.0, offset: float = 0.02) -> Scene:
    rng = np.random.default_rng(seed)
    radius, half_t = 1.0, 0.05
    rb = np.hypot(radius, half_t)
    centres = _place(rng, [rb] * n_bodies, box)
    xs, ns, body = [], [], []
    for bi in range(n_bodies):
        R = _random_rotation(rng)
        p, nrm = _cylinder_surface(rng, pts_per_body, radius, half_t)
        xs.append(p @ R.T + centres[bi])
        ns.append(nrm @ R.T)
        body.append(np.full(pts_per_body, bi))
    x = np.concatenate(xs)
    nx = np.concatenate(ns)
    body = np.concatenate(body)
    y = x - offset * nx
    b = _rigid_rhs(rng, x, nx, body, centres, k)
    return Scene("coins", x, nx, y, b, body, centres,
                 dict(seed=seed, n_bodies=n_bodies, pts_per_body=pts_per_body, offset=offset))


def _fibonacci_sphere(npts):
    kk = np.arange(npts, dtype=np.float64)
    z = 1.0 - (2.0 * kk + 1.0) / npts
    phi = kk * np.pi * (3.0 - np.sqrt(5.0))
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], 1)


def ellipsoids(seed: int = 0, n_bodies: int = 64, pts_per_body: int = 1024, k: int = 8,
               box: float = 20.0, offset_frac: float = 0.1) -> Scene:
    rng = np.random.default_rng(seed)
    axes = rng.uniform(0.5, 1.5, (n_bodies, 3))
    centres = _place(rng, axes.max(axis=1), box)
    unit = _fibonacci_sphere(pts_per_body)
    xs, ns, ys, body = [], [], [], []
    for bi in range(n_bodies):
        a = axes[bi]
        R = _random_rotation(rng)
        p = unit * a
        g = p / (a * a)
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        src = p - offset_frac * a.min() * g
        xs.append(p @ R.T + centres[bi])
        ns.append(g @ R.T)
        ys.append(src @ R.T + centres[bi])
        body.append(np.full(pts_per_body, bi))
    x = np.concatenate(xs)
    nx = np.concatenate(ns)
    y = np.concatenate(ys)
    body = np.concatenate(body)
    b = _rigid_rhs(rng, x, nx, body, centres, k)
    return Scene("ellipsoids", x, nx, y, b, body, centres,
                 dict(seed=seed, n_bodies=n_bodies, pts_per_body=pts_per_body,
                      offset_frac=offset_frac, axes=axes))


def kernel_matrix(x: np.ndarray, nx: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Dense FP64 M[i, j] = -(x_i - y_j) . n_i / (4 pi |x_i - y_j|^3) (small scenes only)."""
    d = x[:, None, :] - y[None, :, :]
    r2 = np.einsum("ijk,ijk->ij", d, d)
    dn = np.einsum("ijk,ik->ij", d, nx)
    return -dn / (FOUR_PI * r2 * np.sqrt(r2))


SCENES = {"coins": coins, "ellipsoids": ellipsoids}
