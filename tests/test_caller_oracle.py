"""Pins of the caller-step oracle (oracle/caller_oracle.py) and of the mesh host
code (fastmbcui_amd/meshes.py) by closed forms, invariants and finite
differences.  CPU only."""
import os

import numpy as np
import pytest
import torch
from scipy import integrate

from fastmbcui_amd import meshes
from oracle import caller_oracle as co

F64 = torch.float64
REF_MESHES = "/root/reference/models/example-robot-data/robots/guesss_description/meshes"


def t(a):
    return torch.as_tensor(np.asarray(a), dtype=F64)


def test_octant_triangle():
    # The flat triangle (e_x, e_y, e_z) covers exactly the first octant from the origin.
    V = t(np.eye(3))
    F = torch.tensor([[0, 1, 2]])
    S = t([[0.0, 0.0, 0.0]])
    assert abs(co.solid_angle(V, F, S).item() - np.pi / 2) < 1e-14
    # reversed orientation flips the sign
    assert abs(co.solid_angle(V, torch.tensor([[0, 2, 1]]), S).item() + np.pi / 2) < 1e-14


@pytest.mark.parametrize("a,d", [(1.0, 0.5), (2.0, 3.0), (0.1, 10.0)])
def test_square_on_axis(a, d):
    h = a / 2
    V = t([[-h, -h, d], [h, -h, d], [h, h, d], [-h, h, d]])
    F = torch.tensor([[0, 1, 2], [0, 2, 3]])
    S = t([[0.0, 0.0, 0.0]])
    got = co.solid_angle(V, F, S).sum().item()
    want = 4 * np.arcsin(a * a / (a * a + 4 * d * d))
    assert abs(got - want) < 1e-13 * max(1.0, abs(want)) + 1e-15


@pytest.mark.parametrize("axes", [(1, 1, 1), (1, 1, 0.3), (1.2, 0.6, 0.6)])
def test_gauss_closed_mesh(axes):
    V, F = meshes.ellipsoid(axes, 2)
    rng = np.random.default_rng(1)
    inside = rng.uniform(-0.25, 0.25, size=(16, 3)) * np.asarray(axes)
    outside = rng.normal(size=(16, 3))
    outside = outside / np.linalg.norm(outside, axis=1, keepdims=True) * 3.0
    Mi = co.solid_angle(t(V), torch.tensor(F), t(inside)).sum(0).numpy()
    Mo = co.solid_angle(t(V), torch.tensor(F), t(outside)).sum(0).numpy()
    assert np.abs(Mi - 4 * np.pi).max() < 1e-11
    assert np.abs(Mo).max() < 1e-11


def test_single_layer_point():
    S = t([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    C = t([[0.0, 3.0, 4.0]])
    SL = co.single_layer(S, C)
    assert abs(SL[0, 0].item() - 0.2) < 1e-15
    assert abs(SL[0, 1].item() - 1 / np.sqrt(26.0)) < 1e-15


def test_oracle_gradients_fd():
    rng = np.random.default_rng(2)
    V, F = meshes.icosphere(0)
    V = t(V + 0.05 * rng.normal(size=V.shape)).requires_grad_()
    S = t(0.3 * rng.normal(size=(5, 3))).requires_grad_()
    Ft = torch.tensor(F)
    assert torch.autograd.gradcheck(lambda v, s: co.solid_angle(v, Ft, s), (V, S))
    C = t(2.0 + rng.normal(size=(4, 3))).requires_grad_()
    X = t(rng.normal(size=(5, 3))).requires_grad_()
    assert torch.autograd.gradcheck(lambda s, c, x: co.single_layer(s, c) @ x, (S, C, X))


def test_icosphere_counts_and_normals():
    for s in range(4):
        V, F = meshes.icosphere(s)
        assert V.shape == (10 * 4 ** s + 2, 3) and F.shape == (20 * 4 ** s, 3)
        # closed, each edge shared by exactly two faces with opposite direction
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        assert len({tuple(x) for x in e}) == len(e)
        assert {tuple(x) for x in e} == {(b, a) for a, b in e}
    V, F = meshes.icosphere(3)
    N = meshes.vertex_normals(V, F)
    assert np.abs(np.linalg.norm(N, axis=1) - 1).max() < 1e-14
    # a sphere's normals are radial, up to the discretisation
    assert np.abs(N - V).max() < 1e-2
    # the 12 icosahedron vertices sit on 5-fold axes: exact there
    assert np.abs(N[:12] - V[:12]).max() < 1e-12
    assert np.allclose(meshes.vertex_normals(V, F), co.vertex_normals(V, F), atol=1e-15)


def test_obj_roundtrip(tmp_path):
    V, F = meshes.ellipsoid((1, 0.5, 0.25), 2)
    p = tmp_path / "m.obj"
    meshes.save_obj(p, V, F)
    V2, F2 = meshes.load_obj(p)
    assert np.array_equal(F, F2) and np.array_equal(V, V2)
    q = tmp_path / "quad.obj"
    q.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1/1 2/2 3/3 4/4\n")
    _, Fq = meshes.load_obj(q)
    assert Fq.tolist() == [[0, 1, 2], [0, 2, 3]]


@pytest.mark.skipif(not os.path.isdir(REF_MESHES), reason="reference meshes not present")
def test_reference_meshes_closed():
    """The shipped OBJ bodies load, and the closed ones give 4 pi from their centroid."""
    for name in sorted(os.listdir(REF_MESHES)):
        if not name.endswith(".obj"):
            continue
        V, F = meshes.load_obj(os.path.join(REF_MESHES, name))
        assert V.shape[0] in (407, 436, 513, 642) and F.shape[0] in (810, 868, 1022, 1280)
        if name.startswith(("ellipse", "sphere")):
            tot = co.solid_angle(t(V), torch.tensor(F), t(V.mean(0, keepdims=True))).sum().item()
            assert abs(abs(tot) - 4 * np.pi) < 1e-9, name


def _spheroid_alpha(a, b, c):
    # Lamb's alpha_0 = a b c int_0^inf dl / ((a^2 + l) sqrt((a^2+l)(b^2+l)(c^2+l)))
    f = lambda l: 1.0 / ((a * a + l) * np.sqrt((a * a + l) * (b * b + l) * (c * c + l)))
    return a * b * c * integrate.quad(f, 0, np.inf, limit=200)[0]


def test_kirchhoff_oracle_added_mass():
    """Translational added mass of an ellipsoid (Lamb, Hydrodynamics sec. 114):
    m_x = rho V alpha_0 / (2 - alpha_0); a sphere gives half the displaced mass.
    Rotational terms of a sphere vanish, and Mf is symmetric."""
    for axes, tol in (((1.0, 1.0, 1.0), 0.005), ((1.0, 1.0, 0.3), 0.01), ((1.2, 0.6, 0.6), 0.01)):
        V, F = meshes.ellipsoid(axes, 3)
        S = meshes.sources(V, F, 0.1)
        Mf, _, _ = co.kirchhoff([t(V)], [torch.tensor(F)], [t(S)], [torch.zeros(3, dtype=F64)],
                                torch.eye(6, dtype=F64), 1.0)
        Mf = Mf.numpy()
        a, b, c = axes
        vol = 4.0 / 3.0 * np.pi * a * b * c
        for i, (p, q, r) in enumerate(((a, b, c), (b, c, a), (c, a, b))):
            al = _spheroid_alpha(p, q, r)
            want = vol * al / (2 - al)
            assert abs(Mf[i, i] - want) < tol * want, (axes, i, Mf[i, i], want)
        assert np.abs(Mf - Mf.T).max() < 1e-9 * np.abs(Mf).max()
        if axes == (1.0, 1.0, 1.0):
            assert np.abs(Mf[3:, 3:]).max() < 1e-5 * Mf[0, 0]   # faceting only


def _kernel_atan_coefficients():
    """Q's coefficients as written in the solid-angle kernel's atan01."""
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "fastmbcui_amd", "csrc",
                            "k_caller.hip")).read()
    body = src[src.index("__device__ inline float atan01"):]
    body = body[:body.index("}")]
    return [np.float32(float(x)) for x in re.findall(r"(-?\d+\.\d+(?:e-?\d+)?)f", body)]


def _atan2_fp32(y, x, q):
    """FP32 emulation of the kernel's atan2_fast (Horner in FP32, fma rounding
    emulated by FP64 products rounded once)."""
    y = y.astype(np.float32)
    x = x.astype(np.float32)
    ax, ay = np.abs(x), np.abs(y)
    mx, mn = np.maximum(ax, ay), np.minimum(ax, ay)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(mx > 0, (mn / mx).astype(np.float32), np.float32(0))
    s = (t * t).astype(np.float32)
    acc = np.full_like(t, q[0])
    for c in q[1:]:
        acc = (acc.astype(np.float64) * s + np.float64(c)).astype(np.float32)
    ts = (t * s).astype(np.float32)
    r = (ts.astype(np.float64) * acc + t).astype(np.float32)
    r = np.where(ay > ax, (np.float32(np.pi / 2) - r).astype(np.float32), r)
    r = np.where(x < 0, (np.float32(np.pi) - r).astype(np.float32), r)
    return np.copysign(r, y)


def test_kernel_atan_polynomial():
    """The kernel's atan coefficients give FP32 atan2 within 2e-7 relative of
    FP64 atan2 over all octants (and atan2(0, 0) = 0)."""
    q = _kernel_atan_coefficients()
    assert len(q) == 8
    rng = np.random.default_rng(0)
    ang = rng.uniform(-np.pi, np.pi, 400000)
    rad = np.exp(rng.uniform(-20, 20, ang.size))
    y, x = rad * np.sin(ang), rad * np.cos(ang)
    y32, x32 = y.astype(np.float32), x.astype(np.float32)
    want = np.arctan2(y32.astype(np.float64), x32.astype(np.float64))
    got = _atan2_fp32(y32, x32, q).astype(np.float64)
    # division by the hardware reciprocal adds about 1 ulp of t
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-30)) < 4e-7
    assert np.max(np.abs(got - want)) < 5e-7
    assert _atan2_fp32(np.zeros(1), np.zeros(1), q)[0] == 0.0
    t = np.linspace(0, 1, 100001)
    full = _atan2_fp32(t, np.ones_like(t), q)
    assert np.max(np.abs(full - np.arctan(t.astype(np.float32).astype(np.float64)))
                  / np.maximum(np.arctan(t), 1e-30)) < 2e-7
