"""GPU parity of the caller steps (face solid angles, single-layer potential)
through the C-ABI and the autograd wrappers, against the FP64 CPU oracle
(oracle/caller_oracle.py) on the same inputs.

Tolerances: FP64 entry points to 1e-11 (absolute on angles up to 2 pi) or
1e-10 normwise; FP32 forward to 4e-6 relative per entry (FP32 arithmetic,
the edge form of the triple product); FP32 backward (computed in FP64) to 1e-6
normwise.  At full size: Gauss's theorem (4 pi inside a closed body, 0
outside) and its gradient (zero: the sum over a closed surface is constant).
"""
import numpy as np
import pytest
import torch

from fastmbcui_amd import _lib, caller, meshes, runtime
from oracle import caller_oracle as co

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)
F64 = torch.float64


def _scene(n_bodies=2, subdiv=3, seed=0, spacing=3.0):
    out = []
    for b in meshes.bodies(n_bodies, subdiv, seed, spacing):
        V = b["V0"] @ b["R"].T + b["y"]
        S = b["S0"] @ b["R"].T + b["y"]
        out.append((V, b["F"], S, b["y"]))
    return out


def _flat(scene):
    """All faces of all bodies against all sources: (V, F, S) with F offset."""
    Vs, Fs, Ss, off = [], [], [], 0
    for V, F, S, _ in scene:
        Vs.append(V)
        Fs.append(F + off)
        Ss.append(S)
        off += V.shape[0]
    return np.concatenate(Vs), np.concatenate(Fs).astype(np.int32), np.concatenate(Ss)


def _gpu(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, dtype)


@pytest.fixture(scope="module")
def ctx():
    return _lib.Context(0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_solid_angle_forward(ctx, dtype):
    V, F, S = _flat(_scene(2, 3))
    got = ctx.solid_angle(_gpu(V, dtype), _gpu(F, torch.int32), _gpu(S, dtype)).cpu().double()
    # oracle on the inputs as the kernel saw them (rounded to dtype)
    Vr, Sr = torch.tensor(V).to(dtype).double(), torch.tensor(S).to(dtype).double()
    want = co.solid_angle(Vr, torch.tensor(F), Sr)
    err = (got - want).abs()
    if dtype == torch.float64:
        assert err.max().item() < 1e-11
    else:
        assert (err <= 4e-6 * want.abs() + 1e-9).all(), err.max().item()
        # the reference's FP32 expression on the same inputs is no more accurate
        ref32 = co.solid_angle(Vr.float(), torch.tensor(F), Sr.float()).double()
        rel = lambda e: (e / want.abs().clamp_min(1e-30)).max().item()
        assert rel(err) <= rel((ref32 - want).abs())


def test_solid_angle_gauss_full_size(ctx):
    """Two bodies of 10,242 vertices / 20,480 faces each (M: 40,960 x 20,484):
    each body's faces sum to 4 pi for its own sources and 0 for the other's."""
    sc = _scene(2, 5, seed=3, spacing=3.0)
    V, F, S = _flat(sc)
    M = ctx.solid_angle(_gpu(V, torch.float32), _gpu(F, torch.int32), _gpu(S, torch.float32))
    nf0, nv0 = sc[0][1].shape[0], sc[0][0].shape[0]
    t0 = M[:nf0].sum(0, dtype=F64).cpu().numpy()
    t1 = M[nf0:].sum(0, dtype=F64).cpu().numpy()
    del M
    assert np.abs(t0[:nv0] - 4 * np.pi).max() < 1e-4 and np.abs(t0[nv0:]).max() < 1e-4
    assert np.abs(t1[nv0:] - 4 * np.pi).max() < 1e-4 and np.abs(t1[:nv0]).max() < 1e-4


def test_solid_angle_errors(ctx):
    V = _gpu(np.eye(3), torch.float32)
    S = _gpu(np.zeros((4, 3)), torch.float32)
    with pytest.raises(_lib.MfsError):
        ctx.solid_angle(V, _gpu(np.array([[0, 1, 3]]), torch.int32), S)
    with pytest.raises(_lib.MfsError):
        ctx.solid_angle_backward(V, _gpu(np.array([[0, -1, 2]]), torch.int32), S,
                                 torch.zeros((1, 4), device=DEV))
    empty = ctx.solid_angle(V, torch.zeros((0, 3), dtype=torch.int32, device=DEV), S)
    assert tuple(empty.shape) == (0, 4)
    # a good call after the errors still works
    M = ctx.solid_angle(V, _gpu(np.array([[0, 1, 2]]), torch.int32), S)
    assert abs(M[0, 0].item() - np.pi / 2) < 1e-6


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_solid_angle_backward(ctx, dtype):
    V, F, S = _flat(_scene(2, 3, seed=1))
    rng = np.random.default_rng(5)
    G = rng.normal(size=(F.shape[0], S.shape[0]))
    gV, gS = ctx.solid_angle_backward(_gpu(V, dtype), _gpu(F, torch.int32), _gpu(S, dtype),
                                      _gpu(G, dtype))
    Vr = torch.tensor(V).to(dtype).double().requires_grad_()
    Sr = torch.tensor(S).to(dtype).double().requires_grad_()
    Gr = torch.tensor(G).to(dtype).double()
    wV, wS = torch.autograd.grad((co.solid_angle(Vr, torch.tensor(F), Sr) * Gr).sum(), (Vr, Sr))
    tol = 1e-10 if dtype == torch.float64 else 1e-6
    for got, want in ((gV, wV), (gS, wS)):
        got = got.cpu().double()
        assert ((got - want).norm() / want.norm()).item() < tol


def test_solid_angle_backward_gauss_full_size(ctx):
    """d/dV and d/dS of the sum over closed surfaces vanish (the sum is 4 pi or 0)."""
    V, F, S = _flat(_scene(2, 4, seed=2))
    G = torch.ones((F.shape[0], S.shape[0]), dtype=F64, device=DEV)
    gV, gS = ctx.solid_angle_backward(_gpu(V, F64), _gpu(F, torch.int32), _gpu(S, F64), G)
    # scale: the gradient of one entry, face 0 seen from the source nearest to it
    c0 = V[F[0]].mean(0)
    s0 = int(np.argmin(np.linalg.norm(S - c0, axis=1)))
    Gone = torch.zeros_like(G)
    Gone[0, s0] = 1.0
    g1v, g1s = ctx.solid_angle_backward(_gpu(V, F64), _gpu(F, torch.int32), _gpu(S, F64), Gone)
    scale = max(g1v.abs().max().item(), g1s.abs().max().item())
    assert scale > 1.0
    assert gV.abs().max().item() < 1e-12 * scale * S.shape[0]
    assert gS.abs().max().item() < 1e-12 * scale * F.shape[0]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("k", [1, 7, 18])
def test_single_layer_apply(ctx, dtype, k):
    V, F, S = _flat(_scene(2, 3, seed=4))
    C = co.face_centers(torch.tensor(V), torch.tensor(F)).numpy()
    X = np.random.default_rng(k).normal(size=(S.shape[0], k))
    got = ctx.single_layer_apply(_gpu(S, dtype), _gpu(C, dtype), _gpu(X, dtype)).cpu().double()
    r = lambda a: torch.tensor(a).to(dtype).double()
    want = co.single_layer(r(S), r(C)) @ r(X)
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    assert ((got - want).norm() / want.norm()).item() < tol
    assert ((got - want).abs().max() / want.abs().max()).item() < 10 * tol


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("k", [3, 18])
def test_single_layer_backward(ctx, dtype, k):
    V, F, S = _flat(_scene(2, 3, seed=6))
    C = co.face_centers(torch.tensor(V), torch.tensor(F)).numpy()
    rng = np.random.default_rng(k)
    X = rng.normal(size=(S.shape[0], k))
    G = rng.normal(size=(C.shape[0], k))
    gS, gC, gX = ctx.single_layer_backward(_gpu(S, dtype), _gpu(C, dtype), _gpu(X, dtype),
                                           _gpu(G, dtype))
    r = lambda a: torch.tensor(a).to(dtype).double().requires_grad_()
    Sr, Cr, Xr = r(S), r(C), r(X)
    wS, wC, wX = torch.autograd.grad(((co.single_layer(Sr, Cr) @ Xr) * r(G).detach()).sum(),
                                     (Sr, Cr, Xr))
    tol = 1e-11 if dtype == torch.float64 else 2e-6
    for got, want in ((gS, wS), (gC, wC), (gX, wX)):
        got = got.cpu().double()
        assert ((got - want).norm() / want.norm()).item() < tol
    # outputs not asked for are skipped
    none = ctx.single_layer_backward(_gpu(S, dtype), _gpu(C, dtype), _gpu(X, dtype), _gpu(G, dtype),
                                     need_s=False, need_c=True, need_x=False)
    assert none[0] is None and none[2] is None
    assert torch.allclose(none[1].cpu().double(), gC.cpu().double(), rtol=0, atol=1e-12 * gC.abs().max().item())


def test_autograd_gradcheck():
    """torch.autograd.gradcheck (FP64 finite differences) of both wrappers on the GPU."""
    rng = np.random.default_rng(7)
    V, F = meshes.icosphere(1)
    Vt = _gpu(V + 0.02 * rng.normal(size=V.shape), F64).requires_grad_()
    Ft = _gpu(F, torch.int32)
    S = _gpu(0.3 * rng.normal(size=(9, 3)), F64).requires_grad_()
    assert torch.autograd.gradcheck(lambda v, s: caller.compute_solid_angle(v, Ft, s), (Vt, S))
    C = _gpu(2.0 + rng.normal(size=(6, 3)), F64).requires_grad_()
    X = _gpu(rng.normal(size=(9, 4)), F64).requires_grad_()
    assert torch.autograd.gradcheck(lambda s, c, x: caller.compute_single_layer(s, c) @ x, (S, C, X))
    with pytest.raises(RuntimeError):
        caller.compute_solid_angle(Vt.detach().cpu(), Ft.cpu(), S.detach().cpu())


def _kirchhoff_inputs(scene, dtype, device):
    V_s = [torch.as_tensor(V).to(device, dtype) for V, _, _, _ in scene]
    F_s = [torch.as_tensor(F).to(device) for _, F, _, _ in scene]
    S_s = [torch.as_tensor(S).to(device, dtype) for _, _, S, _ in scene]
    y_s = [torch.as_tensor(y).to(device, dtype) for _, _, _, y in scene]
    Qp = torch.eye(6 * len(scene), dtype=dtype, device=device)
    return V_s, F_s, S_s, y_s, Qp


def test_kirchhoff_tensor_vs_oracle():
    """Mf of two bodies (KTAB:97-128) against the FP64 oracle with an exact
    least-squares solve, and against Lamb's added mass for a lone sphere."""
    runtime.reset()
    sc = _scene(2, 3, seed=8, spacing=2.6)
    V_s, F_s, S_s, y_s, Qp = _kirchhoff_inputs(sc, torch.float32, DEV)
    Mf, Ks, X = caller.kirchhoff_tensor(V_s, F_s, S_s, y_s, Qp, 1.0, 1.2, 2000, 1e-7)
    oV, oF, oS, oy, oQ = _kirchhoff_inputs(sc, F64, "cpu")
    wMf, _, _ = co.kirchhoff([v.float().double() for v in oV], oF, [s.float().double() for s in oS],
                             oy, oQ, 1.0)
    rel = ((Mf.cpu().double() - wMf).norm() / wMf.norm()).item()
    assert rel < 2e-4, rel
    # a lone unit sphere: half the displaced mass on the translational diagonal
    V, F = meshes.icosphere(4)
    S = meshes.sources(V, F, 0.1)
    one = [(V, F, S, np.zeros(3))]
    V_s, F_s, S_s, y_s, Qp = _kirchhoff_inputs(one, torch.float32, DEV)
    Mf1, _, _ = caller.kirchhoff_tensor(V_s, F_s, S_s, y_s, Qp, 1.0, 1.2, 2000, 1e-7)
    d = Mf1.diagonal()[:3].cpu().double().numpy()
    assert np.abs(d - 2 * np.pi / 3).max() < 0.003 * 2 * np.pi / 3, d


def test_kirchhoff_gradient_vs_oracle():
    """dL/dy for L = 0.5 v^T Mf v with two bodies translated by y (the chain
    of L.backward(), KTAB:136-139): GPU autograd against the oracle's."""
    runtime.reset()
    sc = _scene(2, 2, seed=9, spacing=2.4)
    v = torch.tensor([0.3, -0.2, 0.5, 0.1, 0.2, -0.1, -0.4, 0.1, 0.2, 0.0, 0.3, 0.1], dtype=F64)

    def run(dtype, device, oracle):
        V_s, F_s, S_s, y_s, Qp = _kirchhoff_inputs(sc, dtype, device)
        shift = torch.zeros((2, 3), dtype=dtype, device=device, requires_grad=True)
        V_s = [V + shift[i] for i, V in enumerate(V_s)]
        S_s = [S + shift[i] for i, S in enumerate(S_s)]
        y_s = [y + shift[i] for i, y in enumerate(y_s)]
        if oracle:
            Mf = co.kirchhoff(V_s, F_s, S_s, y_s, Qp, 1.0)[0]
        else:
            Mf = caller.kirchhoff_tensor(V_s, F_s, S_s, y_s, Qp, 1.0, 1.2, 4000, 1e-8)[0]
        vv = v.to(device, dtype)
        L = 0.5 * vv @ Mf @ vv
        L.backward()
        return shift.grad.cpu().double()

    got = run(torch.float32, DEV, False)
    want = run(F64, "cpu", True)
    rel = ((got - want).norm() / want.norm()).item()
    assert rel < 5e-3, (rel, got, want)


def test_kirchhoff_step_vs_oracle():
    """The whole caller step (KTAB:88-163, no drag) for the coin driver's model
    (two PlanarYZV2 links): Mf, dL/dq and d(Mf_KDetached)/dq against the FP64
    oracle pipeline with an exact least-squares solve and the entry-by-entry
    backward loop."""
    from fastmbcui_amd import kinematics as kin
    runtime.reset()
    table = kin.JointTable([kin.Joint(kind=kin.KINDS["PlanarYZV2"]),
                            kin.Joint(kind=kin.KINDS["PlanarYZV2"])])
    shapes = [meshes.ellipsoid((1.0, 1.0, 0.3), 2), meshes.ellipsoid((1.0, 0.6, 0.6), 2)]
    bodies64 = [(torch.tensor(V), torch.tensor(F), torch.tensor(meshes.sources(V, F, 0.1)))
                for V, F in shapes]
    bodies = [(V.to(DEV, torch.float32), F.to(DEV), S.to(DEV, torch.float32)) for V, F, S in bodies64]
    q = np.array([0.1, -0.2, 0.3, 2.4, 0.4, -0.5])
    v = np.array([0.3, -0.2, 0.5, 0.1, 0.2, -0.4])
    out = caller.kirchhoff_step(table, bodies, q, v, 1.0, 1.2, 4000, 1e-8)

    qt = torch.tensor(q, requires_grad=True)
    R_s, y_s, Qp_s = kin.poses(table, qt)
    Qp = torch.cat(Qp_s, 0)
    r32 = lambda t: t.float().double()     # the inputs as the GPU saw them
    V_s = [r32(V) @ R.T + y for (V, _, _), R, y in zip(bodies64, R_s, y_s)]
    S_s = [r32(S) @ R.T + y for (_, _, S), R, y in zip(bodies64, R_s, y_s)]
    Mf, Ks, _ = co.kirchhoff(V_s, [F for _, F, _ in bodies64], S_s, y_s, Qp, 1.0)
    vt = torch.tensor(v)
    (gq,) = torch.autograd.grad(0.5 * vt @ Mf @ vt, qt, retain_graph=True)
    R6n = kin.r6n(R_s)
    K = (R6n.T @ co.iq_matrix(2) @ Ks).detach()
    J = co.jacobian_loop(Qp.T @ R6n @ K, qt)

    rel = lambda a, b: ((a.cpu().double() - b).norm() / b.norm()).item()
    assert rel(out["Mf"], Mf.detach()) < 2e-4
    assert rel(out["K"], K) < 2e-4
    assert rel(out["grad_L_q"], gq) < 5e-3, (out["grad_L_q"], gq)
    assert rel(out["grad_Mf_q_KDetached"], J) < 2e-4
