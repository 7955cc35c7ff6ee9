"""Scene I/O: URDF reading, URDF rotations, the shipped problems' joint tables
and bodies (CPU); one GPU step per shipped problem at the end."""
import os

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from fastmbcui_amd import kinematics as kin
from fastmbcui_amd import problems
from oracle import caller_oracle as co

REF_MODELS = "/root/reference/models"

URDF = """<?xml version="1.0" ?>
<robot name="t">
  <link name="a">
    <visual><origin rpy="0.1 -0.2 0.3" xyz="1 2 3"/>
      <geometry><mesh filename="package://x/a.obj" scale="2 2 2"/></geometry></visual>
  </link>
  <link name="empty"/>
  <link name="b">
    <visual><geometry><mesh filename="package://x/b.obj"/></geometry></visual>
  </link>
  <joint name="j" type="revolute"><parent link="a"/><child link="b"/>
    <origin xyz="0 0 1" rpy="0 0 0"/><axis xyz="0 1 0"/></joint>
</robot>
"""


def test_read_urdf(tmp_path):
    p = tmp_path / "t.urdf"
    p.write_text(URDF)
    vis, joints = problems.read_urdf(p)
    assert [v.link for v in vis] == ["a", "b"]
    assert vis[0].mesh == "package://x/a.obj" and np.allclose(vis[0].scale, 2)
    assert np.allclose(vis[0].xyz, [1, 2, 3]) and np.allclose(vis[0].rpy, [0.1, -0.2, 0.3])
    assert np.allclose(vis[1].xyz, 0) and np.allclose(vis[1].scale, 1)
    assert joints[0]["type"] == "revolute" and joints[0]["parent"] == "a" and joints[0]["child"] == "b"
    assert np.allclose(joints[0]["axis"], [0, 1, 0])


def test_rpy_matches_extrinsic_xyz():
    rng = np.random.default_rng(0)
    for _ in range(20):
        rpy = rng.uniform(-np.pi, np.pi, 3)
        assert np.allclose(problems.rpy_matrix(rpy), Rotation.from_euler("xyz", rpy).as_matrix(),
                           atol=1e-14)


def _check(P):
    t = P["table"]
    assert t.nq == len(P["q_init"]) and len(P["bodies"]) == len(t.joints)
    R_s, y_s, Qp_s = kin.poses(t, torch.tensor(P["q_init"]))
    assert all(Q.shape == (6, t.nv) for Q in Qp_s)
    for V, F, S in P["bodies"]:
        tot = co.solid_angle(torch.tensor(V), torch.tensor(F), torch.tensor(S)).sum(0).numpy()
        assert np.abs(tot - 4 * np.pi).max() < 1e-8   # every source inside its closed body


@pytest.mark.parametrize("name", sorted(problems.PROBLEMS))
def test_stand_in_problems(name):
    _check(problems.load_problem(name))


@pytest.mark.skipif(not os.path.isdir(REF_MODELS), reason="reference models not present")
@pytest.mark.parametrize("name,counts", [("coinsPlanar_v5", [642, 642]),
                                         ("grasp_v7x10", [436, 407, 407, 642]),
                                         ("leaf_v1x10", [642, 642, 513])])
def test_shipped_problems_from_assets(name, counts):
    P = problems.load_problem(name, REF_MODELS)
    assert [b[0].shape[0] for b in P["bodies"]] == counts
    _check(P)


def test_unknown_problem():
    with pytest.raises(ValueError):
        problems.load_problem("nope")


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(problems.PROBLEMS))
def test_shipped_problem_step_on_device(name):
    """One Kirchhoff step per shipped problem (stand-in bodies) at the driver's
    q_init and solver settings: Mf symmetric positive semidefinite (an added
    mass), gradients finite."""
    from fastmbcui_amd import caller, runtime
    runtime.reset()
    dev = torch.device("cuda", 0)
    P = problems.load_problem(name)
    bodies = [(torch.tensor(V, dtype=torch.float32, device=dev), torch.tensor(F, device=dev),
               torch.tensor(S, dtype=torch.float32, device=dev)) for V, F, S in P["bodies"]]
    v = np.linspace(-1, 1, P["table"].nv)
    out = caller.kirchhoff_step(P["table"], bodies, P["q_init"], v, P["rho_fluid"], P["rho_mfs"],
                                P["pcg_max_iter"], P["tol_r"])
    Mf = out["Mf"].double().cpu()
    scale = Mf.abs().max().item()
    assert scale > 0 and torch.isfinite(Mf).all()
    assert (Mf - Mf.T).abs().max().item() < 2e-2 * scale
    assert torch.linalg.eigvalsh(0.5 * (Mf + Mf.T)).min().item() > -2e-2 * scale
    assert torch.isfinite(out["grad_L_q"]).all() and torch.isfinite(out["grad_Mf_q_KDetached"]).all()
