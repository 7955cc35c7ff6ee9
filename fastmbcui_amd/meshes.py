"""Triangle meshes and MFS sources for the caller steps.

The reference reads its bodies from OBJ files with meshio
(core/KTAB_Mfs_spclf1d.py:38-41; models/example-robot-data/robots/
guesss_description/meshes/*.obj: meshio-written OBJ, `v x y z` and
`f i j k` lines, 1-based), computes angle-weighted vertex normals once
(utils/torch_utils.py:4-47) and places one source per vertex, moved inward
along its normal by the body's offset (KTAB:50-52).  This module does the
same on the host with numpy: an OBJ reader, icosphere / ellipsoid meshes of
the same family as the shipped ones (the 642-vertex, 1280-face spheres and
ellipses are icospheres of subdivision 3), the vertex normals and the sources.
"""
from __future__ import annotations

import numpy as np


def load_obj(path):
    """(V float64 nv x 3, F int32 nf x 3) from an OBJ file; polygons become fans."""
    verts, faces = [], []
    with open(path, "r") as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f":
                idx = []
                for tok in parts[1:]:
                    i = int(tok.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(verts) + i)
                for t in range(1, len(idx) - 1):
                    faces.append([idx[0], idx[t], idx[t + 1]])
    V = np.asarray(verts, np.float64).reshape(-1, 3)
    F = np.asarray(faces, np.int32).reshape(-1, 3)
    if F.size and (F.min() < 0 or F.max() >= V.shape[0]):
        raise ValueError(f"{path}: face index out of range")
    return V, F


def save_obj(path, V, F):
    with open(path, "w") as fh:
        for v in np.asarray(V, np.float64):
            fh.write(f"v {v[0]:.17g} {v[1]:.17g} {v[2]:.17g}\n")
        for f in np.asarray(F, np.int64):
            fh.write(f"f {f[0] + 1} {f[1] + 1} {f[2] + 1}\n")


def icosphere(subdiv=3):
    """Unit icosphere: 10 * 4^subdiv + 2 vertices, 20 * 4^subdiv faces, outward
    counter-clockwise faces."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
         [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
         [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
         [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    V = [np.asarray(v, np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(subdiv):
        mid = {}

        def midpoint(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in mid:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                mid[key] = len(V) - 1
            return mid[key]

        nxt = []
        for a, b, c in F:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nxt += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        F = nxt
    return np.asarray(V, np.float64), np.asarray(F, np.int32)


def ellipsoid(axes=(1.0, 1.0, 1.0), subdiv=3):
    V, F = icosphere(subdiv)
    return V * np.asarray(axes, np.float64), F


def vertex_normals(V, F):
    """Angle-weighted vertex normals (the reference's compute_normal_vertex,
    utils/torch_utils.py:4-47): face normals weighted by the face's angle at
    the vertex, summed and normalised."""
    V = np.asarray(V, np.float64)
    F = np.asarray(F, np.int64)
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e01, e02, e12 = p1 - p0, p2 - p0, p2 - p1
    fn = np.cross(e01, e02)
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)

    def angle(a, b):
        c = np.einsum("ij,ij->i", a, b) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))
        return np.arccos(np.clip(c, -1.0, 1.0))

    ang0 = angle(e01, e02)
    ang1 = angle(-e01, e12)
    ang2 = np.pi - ang0 - ang1
    acc = np.zeros_like(V)
    for i, a in enumerate((ang0, ang1, ang2)):
        np.add.at(acc, F[:, i], fn * a[:, None])
    return acc / np.linalg.norm(acc, axis=1, keepdims=True)


def sources(V, F, offset):
    """One source per vertex, moved inward by `offset` along the vertex normal
    (KTAB:50-52)."""
    return np.asarray(V, np.float64) - float(offset) * vertex_normals(V, F)


def rotation(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def bodies(n_bodies=3, subdiv=3, seed=0, spacing=3.0, offset=0.1):
    """Body-frame meshes of a small free-floating scene: list of dicts with V0,
    F, S0 (sources) and a pose (R, y).  Shapes alternate between a sphere, a
    flat ellipsoid ("coin") and a long ellipsoid, as in the shipped scenes."""
    rng = np.random.default_rng(seed)
    shapes = [(1.0, 1.0, 1.0), (1.0, 1.0, 0.3), (1.2, 0.6, 0.6)]
    out = []
    for b in range(n_bodies):
        V0, F = ellipsoid(shapes[b % len(shapes)], subdiv)
        R = rotation(rng.normal(size=3), rng.uniform(0, np.pi))
        y = np.array([spacing * b, 0.3 * rng.normal(), 0.3 * rng.normal()])
        out.append(dict(V0=V0, F=F, S0=sources(V0, F, offset), R=R, y=y))
    return out
