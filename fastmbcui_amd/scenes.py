"""Seeded synthetic MFS scenes (collocation points, normals, sources, right-hand sides).

The shipped drivers of fastMBCUI build their scenes from OBJ meshes
(core/KTAB_Mfs_spclf1d.py:33-51).  The workloads measured here are point
scenes instead:

* ``coins(seed)``: 8 cylinders (radius 1, thickness 0.1), 512 area-uniform
  surface points each (N = 4096), sources at x - 0.02 n.
* ``ellipsoids(seed)``: 64 ellipsoids with semi-axes ~ U(0.5, 1.5), 1024
  Fibonacci-mapped points each (N = 65,536), sources at x - 0.1 min(a,b,c) n.
* ``mixed_bodies(seed)``: 128 bodies (coins, ellipsoids, rounded boxes with
  edge radius 0.1), 2048 points each (N = 262,144), in rigid motion at
  constant v, w ~ N(0, 0.5^2); ``RigidScene.advance(dt)`` steps it.

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


def _rounded_box_surface(rng, npts, half, r):
    """Area-uniform samples on a box with half-extents `half` whose edges and
    corners are rounded with radius r (the inner box grown by a ball of radius
    r), body frame.  Returns points and outward unit normals."""
    inner = np.asarray(half, dtype=np.float64) - r
    ex, ey, ez = 2 * inner
    faces = [ey * ez, ey * ez, ex * ez, ex * ez, ex * ey, ex * ey]
    edges = [2 * inner[d] * np.pi * r / 2 for d in (0, 1, 2) for _ in range(4)]
    corner = 4 * np.pi * r * r
    areas = np.array(faces + edges + [corner])
    part = rng.choice(len(areas), size=npts, p=areas / areas.sum())
    nrm = np.zeros((npts, 3))
    base = np.zeros((npts, 3))   # closest point on the inner box
    for f in range(6):           # faces: +x, -x, +y, -y, +z, -z
        sel = part == f
        d, sgn = f // 2, (1.0 if f % 2 == 0 else -1.0)
        o1, o2 = [a for a in (0, 1, 2) if a != d]
        base[sel, d] = sgn * inner[d]
        base[sel, o1] = rng.uniform(-inner[o1], inner[o1], sel.sum())
        base[sel, o2] = rng.uniform(-inner[o2], inner[o2], sel.sum())
        nrm[sel, d] = sgn
    for e in range(12):          # edges along axis d, quadrant q of the other two
        sel = part == 6 + e
        d, q = e // 4, e % 4
        o1, o2 = [a for a in (0, 1, 2) if a != d]
        s1, s2 = (1.0 if q & 1 else -1.0), (1.0 if q & 2 else -1.0)
        th = rng.uniform(0.0, np.pi / 2, sel.sum())
        base[sel, d] = rng.uniform(-inner[d], inner[d], sel.sum())
        base[sel, o1] = s1 * inner[o1]
        base[sel, o2] = s2 * inner[o2]
        nrm[sel, o1] = s1 * np.cos(th)
        nrm[sel, o2] = s2 * np.sin(th)
    sel = part == 18             # corners: a uniform direction picks the octant
    g = rng.standard_normal((sel.sum(), 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    base[sel] = np.sign(g) * inner
    nrm[sel] = g
    return base + r * nrm, nrm


def _rotation_step(w, dt):
    """exp(dt [w]_x) by Rodrigues' formula, one 3x3 per row of w."""
    th = np.linalg.norm(w, axis=1) * dt
    out = np.repeat(np.eye(3)[None], w.shape[0], axis=0)
    nz = th > 0
    a = w[nz] / np.linalg.norm(w[nz], axis=1, keepdims=True)
    K = np.zeros((a.shape[0], 3, 3))
    K[:, 0, 1], K[:, 0, 2], K[:, 1, 2] = -a[:, 2], a[:, 1], -a[:, 0]
    K[:, 1, 0], K[:, 2, 0], K[:, 2, 1] = a[:, 2], -a[:, 1], a[:, 0]
    s, c = np.sin(th[nz])[:, None, None], np.cos(th[nz])[:, None, None]
    out[nz] = np.eye(3)[None] + s * K + (1 - c) * (K @ K)
    return out


@dataclass
class RigidScene:
    """Bodies in rigid motion at constant velocities (v, w per body).

    Body-frame points/normals/sources are fixed; `advance(dt)` moves the centres
    by v dt and turns the frames by exp(dt [w]_x); `scene()` gives the current
    world-frame Scene whose single right-hand side is the Neumann data of the
    motion, n . (v + w x (x - c))."""

    name: str
    local_x: np.ndarray
    local_n: np.ndarray
    local_y: np.ndarray
    body_of_point: np.ndarray
    rot: np.ndarray       # n_bodies x 3 x 3
    centres: np.ndarray   # n_bodies x 3
    v: np.ndarray
    w: np.ndarray
    meta: dict = field(default_factory=dict)

    def advance(self, dt: float) -> None:
        self.centres = self.centres + dt * self.v
        self.rot = _rotation_step(self.w, dt) @ self.rot

    def scene(self) -> Scene:
        R = self.rot[self.body_of_point]
        c = self.centres[self.body_of_point]
        x = np.einsum("nij,nj->ni", R, self.local_x) + c
        nx = np.einsum("nij,nj->ni", R, self.local_n)
        y = np.einsum("nij,nj->ni", R, self.local_y) + c
        vel = self.v[self.body_of_point] + np.cross(self.w[self.body_of_point], x - c)
        b = np.einsum("ij,ij->i", nx, vel)[:, None]
        return Scene(self.name, x, nx, y, b, self.body_of_point, self.centres.copy(), dict(self.meta))


def mixed_bodies(seed: int = 0, n_bodies: int = 128, pts_per_body: int = 2048, box: float = 32.0,
                 speed: float = 0.5, edge_radius: float = 0.1, kinds=(0, 1, 2)) -> RigidScene:
    """Coins (radius 1, thickness 0.1, sources 0.02 inside), ellipsoids
    (semi-axes ~ U(0.5, 1.5), Fibonacci points, sources 0.1 min axis inside) and
    rounded boxes (half-extents ~ U(0.5, 1.5), edge radius 0.1, sources 0.05
    inside) -- kinds 0, 1, 2, taken in turn from `kinds` by body index (a third
    of each by default); v, w ~ N(0, speed^2)."""
    rng = np.random.default_rng(seed)
    shape = np.asarray(kinds)[np.arange(n_bodies) % len(kinds)]
    dims = rng.uniform(0.5, 1.5, (n_bodies, 3))
    radii = np.where(shape == 0, np.hypot(1.0, 0.05),
                     np.where(shape == 1, dims.max(axis=1), np.linalg.norm(dims, axis=1)))
    centres = _place(rng, radii, box)
    unit = _fibonacci_sphere(pts_per_body)
    lx, ln, ly, body, rots = [], [], [], [], []
    for bi in range(n_bodies):
        rots.append(_random_rotation(rng))
        if shape[bi] == 0:
            p, nrm = _cylinder_surface(rng, pts_per_body, 1.0, 0.05)
            src = p - 0.02 * nrm
        elif shape[bi] == 1:
            a = dims[bi]
            p = unit * a
            nrm = p / (a * a)
            nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
            src = p - 0.1 * a.min() * nrm
        else:
            p, nrm = _rounded_box_surface(rng, pts_per_body, dims[bi], edge_radius)
            src = p - 0.5 * edge_radius * nrm
        lx.append(p)
        ln.append(nrm)
        ly.append(src)
        body.append(np.full(pts_per_body, bi))
    v = speed * rng.standard_normal((n_bodies, 3))
    w = speed * rng.standard_normal((n_bodies, 3))
    return RigidScene("mixed", np.concatenate(lx), np.concatenate(ln), np.concatenate(ly),
                      np.concatenate(body), np.array(rots), centres, v, w,
                      dict(seed=seed, n_bodies=n_bodies, pts_per_body=pts_per_body, box=box,
                           speed=speed, shapes=shape))


def batch_scene(seed: int, bodies=(2, 32), points=(256, 2048)) -> Scene:
    """One scene of the uneven batch: body count and points per body drawn
    log-uniform from the given ranges (seeded), coins and ellipsoids in turn,
    one right-hand side of a rigid motion with v, w ~ N(0, 1) per body."""
    rng = np.random.default_rng(seed)
    nb = int(round(np.exp(rng.uniform(np.log(bodies[0]), np.log(bodies[1])))))
    ppb = int(round(np.exp(rng.uniform(np.log(points[0]), np.log(points[1])))))
    box = 3.5 * nb ** (1.0 / 3.0) + 3.0
    rs = mixed_bodies(seed=seed + 1_000_003, n_bodies=nb, pts_per_body=ppb, box=box, speed=1.0,
                      kinds=(0, 1))
    sc = rs.scene()
    sc.name = f"batch{seed}"
    sc.meta.update(n_bodies=nb, pts_per_body=ppb)
    return sc


def kernel_matrix(x: np.ndarray, nx: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Dense FP64 M[i, j] = -(x_i - y_j) . n_i / (4 pi |x_i - y_j|^3) (small scenes only)."""
    d = x[:, None, :] - y[None, :, :]
    r2 = np.einsum("ijk,ijk->ij", d, d)
    dn = np.einsum("ijk,ik->ij", d, nx)
    return -dn / (FOUR_PI * r2 * np.sqrt(r2))


SCENES = {"coins": coins, "ellipsoids": ellipsoids, "mixed": mixed_bodies}
