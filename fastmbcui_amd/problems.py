"""The shipped scenes as joint tables and link meshes (no pinocchio, no meshio).

fastMBCUI builds its three problems in models/utils/makeModels.py
(makeModel_coinsPlanar :17-45, makeModel_grasp :47-80, makeModel_leaf :82-101)
as pinocchio models with hand-made joints, loads each link's visual mesh from
the problem's URDF (utils/pin_utils.py:5-13, pinocchio.buildModelsFromUrdf)
and places it by the visual origin (core/KTAB_Mfs_spclf1d.py:34-41); the
drivers give the source offsets and the start configuration
(simuspclf1d_mfs_coins.py:12-27, simuspclf1d_btPin_mfs_grasp.py:16-34,
simuspclf1d_btPin_mfs_leaf_floor.py:17-36).  Here:

* `read_urdf` reads a URDF's links (visual mesh file, scale, origin) and joints
  with xml.etree;
* `load_problem(name, models_dir)` returns the joint table of the matching
  makeModel_* builder, one body per link (V0, F, S0 in the link frame: the OBJ
  vertices moved by the visual origin, sources moved inward by the driver's
  offset along the vertex normals, KTAB:50-52), and the driver's q_init and
  solver settings.

Without `models_dir` (the reference's `models/` folder), `load_problem`
builds stand-in bodies: 642-vertex ellipsoids of each mesh's extents (the
meshes are 407-642-vertex closed surfaces), so the scenes' sizes and
kinematics can be exercised where the asset files are absent.
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass

import numpy as np

from . import kinematics as kin
from . import meshes


def rpy_matrix(rpy):
    """URDF roll-pitch-yaw: R = Rz(yaw) Ry(pitch) Rx(roll)."""
    r, p, y = rpy
    cr, sr, cp, sp, cy, sy = np.cos(r), np.sin(r), np.cos(p), np.sin(p), np.cos(y), np.sin(y)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


@dataclass
class UrdfVisual:
    link: str
    mesh: str          # path as written (package://...)
    scale: np.ndarray
    xyz: np.ndarray
    rpy: np.ndarray


def _floats(text, default):
    return np.array([float(x) for x in text.split()]) if text else np.array(default, np.float64)


def read_urdf(path):
    """(visuals in document order, joints as dicts)."""
    root = ET.parse(path).getroot()
    visuals, joints = [], []
    for link in root.findall("link"):
        vis = link.find("visual")
        if vis is None:
            continue
        mesh = vis.find("geometry/mesh")
        if mesh is None:
            continue
        org = vis.find("origin")
        visuals.append(UrdfVisual(
            link=link.get("name"), mesh=mesh.get("filename"),
            scale=_floats(mesh.get("scale"), [1, 1, 1]),
            xyz=_floats(org.get("xyz") if org is not None else None, [0, 0, 0]),
            rpy=_floats(org.get("rpy") if org is not None else None, [0, 0, 0])))
    for j in root.findall("joint"):
        org = j.find("origin")
        axis = j.find("axis")
        joints.append(dict(name=j.get("name"), type=j.get("type"),
                           parent=j.find("parent").get("link"), child=j.find("child").get("link"),
                           xyz=_floats(org.get("xyz") if org is not None else None, [0, 0, 0]),
                           rpy=_floats(org.get("rpy") if org is not None else None, [0, 0, 0]),
                           axis=_floats(axis.get("xyz") if axis is not None else None, [1, 0, 0])))
    return visuals, joints


def _resolve(mesh_path, models_dir):
    prefix = "package://"
    rel = mesh_path[len(prefix):] if mesh_path.startswith(prefix) else mesh_path
    return os.path.join(models_dir, rel)


# Joint tables of makeModels.py (parents as joint indices, -1 = world) and the
# drivers' settings.  grasp's finger joints sit at +-offset_ = 0 along z.
_K = kin.KINDS
PROBLEMS = {
    "coinsPlanar_v5": dict(
        urdf="coinsPlanar_v1", joints=[(_K["PlanarYZV2"], -1, 0.0), (_K["PlanarYZV2"], -1, 0.0)],
        offsets=[0.0012, 0.0012], q_init=[0.0, 0, -np.pi / 10, 0.02, -0.03, -np.pi / 4],
        rho_mfs=1.2, pcg_max_iter=75, tol_r=2.7e-3, rho_fluid=998.0,
        stand_in=[((0.02, 0.003, 0.02), (0, 0, 0)), ((0.02, 0.003, 0.02), (0, 0, 0))]),
    "grasp_v7x10": dict(
        urdf="grasp_v7x10", joints=[(_K["PY"], -1, 0.0), (_K["PZ"], 0, 0.0), (_K["PZ"], 0, -0.0),
                                    (_K["PlanarYZV2"], -1, 0.0)],
        offsets=[0.0012 * 20, 0.0012 * 60, 0.0012 * 60, 0.0012 * 10],
        q_init=[0.0, -0.4, 0.4, 1.4, 0, 0.4],
        rho_mfs=1.15, pcg_max_iter=75, tol_r=2.7e-3, rho_fluid=998.0,
        stand_in=[((0.06, 0.06, 0.9), (0, -0.06, 0)), ((0.4, 0.8, 0.24), (0, 0.8, -0.24)),
                  ((0.4, 0.8, 0.24), (0, 0.8, 0.24)), ((0.2, 0.05, 0.2), (0, 0, 0))]),
    "leaf_v1x10": dict(
        urdf="leaf_v1x10", joints=[(_K["PZ"], -1, 0.0), (_K["PlanarYZV2"], -1, 0.0), (_K["PZ"], -1, 0.0)],
        offsets=[0.0012 * 30, 0.0012 * 10, 0.21], q_init=[0.875, 0.7, -0.875, 1.57079633, -1.48],
        rho_mfs=1.15, pcg_max_iter=75, tol_r=2.7e-3, rho_fluid=998.0,
        stand_in=[((0.6, 0.6, 0.6), (0, 0, 0)), ((0.2, 0.03, 0.2), (0, 0, 0)),
                  ((5.0, 5.0, 0.5), (0, 0, 0))]),
}


def load_problem(name, models_dir=None):
    """dict(table, bodies=[(V0, F, S0)], q_init, offsets, rho_mfs, pcg_max_iter,
    tol_r, rho_fluid, names) for one of PROBLEMS.  Arrays are numpy FP64 (F
    int32); bodies are in their link frames."""
    if name not in PROBLEMS:
        raise ValueError(f"unknown problem {name!r}; known: {sorted(PROBLEMS)}")
    P = PROBLEMS[name]
    table = kin.JointTable([kin.Joint(kind=k, parent=p, r=np.array([0.0, 0.0, z]))
                            for k, p, z in P["joints"]])
    bodies, names = [], []
    if models_dir is not None:
        urdf = os.path.join(models_dir, "example-robot-data", "robots", "guesss_description",
                            "robots", P["urdf"] + ".urdf")
        visuals, _ = read_urdf(urdf)
        if len(visuals) != len(P["joints"]):
            raise ValueError(f"{urdf}: {len(visuals)} visual links for {len(P['joints'])} joints")
        for vis, off in zip(visuals, P["offsets"]):
            V, F = meshes.load_obj(_resolve(vis.mesh, models_dir))
            V = (V * vis.scale) @ rpy_matrix(vis.rpy).T + vis.xyz
            bodies.append((V, F, meshes.sources(V, F, off)))
            names.append(vis.link)
    else:
        for b, ((axes, centre), off) in enumerate(zip(P["stand_in"], P["offsets"])):
            V, F = meshes.ellipsoid(axes, 3)
            V = V + np.asarray(centre, np.float64)
            # sources at the driver's offset, capped at 40% of the thinnest semi-axis
            bodies.append((V, F, meshes.sources(V, F, min(off, 0.4 * min(axes)))))
            names.append(f"stand-in {b}")
    return dict(table=table, bodies=bodies, q_init=np.array(P["q_init"], np.float64),
                offsets=list(P["offsets"]), rho_mfs=P["rho_mfs"], pcg_max_iter=P["pcg_max_iter"],
                tol_r=P["tol_r"], rho_fluid=P["rho_fluid"], names=names)
