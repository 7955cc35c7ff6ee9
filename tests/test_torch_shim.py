"""libmfs_torch.so, the operator library the reference's op package loads
(core/mfs_torch_py/__init__.py:7-8 -> ./build/mfs_torch/csrc/libmfs_torch.so):
its two schemas (mfs_torch_ops.cpp:8-9), the registrations the reference's
ops.py makes on top of it (register_fake, register_autograd), and -- on the
GPU -- its CUDA-key implementations against the oracle.

Each check runs in a fresh interpreter: the library defines the mfs_torch
namespace, which fastmbcui_amd.mfs_torch_py (imported by other tests) also
defines, and a namespace is defined once per process."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "build", "mfs_torch", "csrc", "libmfs_torch.so")

LOAD = f"""
import sys, json
sys.path.insert(0, {ROOT!r})
import torch
torch.ops.load_library({SHIM!r})
"""

# What the reference's ops.py registers on top of the loaded library: fake
# kernels returning (n, k) float tensors and the backward through VpL.
REGISTER = """
@torch.library.register_fake("mfs_torch::MinvF_float_fastGrad")
def _(M, F, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose=False, solve_verbose=False):
    return torch.empty([Vs_s.shape[0], F.shape[1]], dtype=torch.float, device=M.device)

@torch.library.register_fake("mfs_torch::VpL_fast_float")
def _(M, pL, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose=False, solve_verbose=False):
    return torch.empty([Vs_s.shape[0], pL.shape[1]], dtype=torch.float, device=M.device)

def _setup(ctx, inputs, output):
    M, F, Vs_s, rho, it, tol, tv, sv = inputs
    ctx.save_for_backward(M, F, Vs_s, output)
    ctx.args = (rho, it, tol, tv, sv)

def _backward(ctx, g):
    M, F, Vs_s, A = ctx.saved_tensors
    V = torch.ops.mfs_torch.VpL_fast_float(M, g.contiguous(), Vs_s, *ctx.args)
    MV = M @ V
    return (F - M @ A) @ V.T - MV @ A.T, MV, None, None, None, None, None, None

torch.library.register_autograd("mfs_torch::MinvF_float_fastGrad", _backward, setup_context=_setup)
"""


def _run(code, timeout=300):
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_shim_is_built():
    assert os.path.exists(SHIM), "build it with python -m fastmbcui_amd.build"


def test_shim_defines_the_reference_schemas_and_takes_its_registrations():
    out = _run(LOAD + REGISTER + """
from torch._subclasses.fake_tensor import FakeTensorMode
s = {n: str(getattr(torch.ops.mfs_torch, n).default._schema)
     for n in ("VpL_fast_float", "MinvF_float_fastGrad")}
with FakeTensorMode():
    o = torch.ops.mfs_torch.MinvF_float_fastGrad(torch.empty(10, 7), torch.empty(10, 3), torch.empty(7, 3),
                                                 3.0, 10, 1e-6)
    s["fake_shape"] = list(o.shape)
print(json.dumps(s))
""")
    got = json.loads(out.strip().splitlines()[-1])
    assert got["VpL_fast_float"] == ("mfs_torch::VpL_fast_float(Tensor M, Tensor pL, Tensor Vs_s, float rho, "
                                     "int pcg_max_iter, float tol_r, bool timing_verbose=False, "
                                     "bool solve_verbose=False) -> Tensor")
    assert got["MinvF_float_fastGrad"] == ("mfs_torch::MinvF_float_fastGrad(Tensor M, Tensor F, Tensor Vs_s, "
                                           "float rho, int pcg_max_iter, float tol_r, "
                                           "bool timing_verbose=False, bool solve_verbose=False) -> Tensor")
    assert got["fake_shape"] == [7, 3]


@pytest.mark.gpu
def test_shim_solves_on_the_gpu_like_the_oracle():
    out = _run(LOAD + REGISTER + """
import numpy as np
import oracle
from fastmbcui_amd import scenes
sc = scenes.coins(seed=0, n_bodies=3, pts_per_body=160, k=2, box=6.0)
M = scenes.kernel_matrix(sc.x, sc.nx, sc.y).astype(np.float32)
F = sc.b.astype(np.float32)
Vs = torch.from_numpy(sc.y.astype(np.float32))
dev = torch.device("cuda", 0)
res = {}
try:
    torch.ops.mfs_torch.VpL_fast_float(torch.from_numpy(M).to(dev), torch.ones(sc.n, 1, device=dev), Vs,
                                       3.0, 50, 1e-6)
except RuntimeError as e:
    res["vpl_first"] = str(e)
Md = torch.from_numpy(M).to(dev).requires_grad_()
Fd = torch.from_numpy(F).to(dev).requires_grad_()
X = torch.ops.mfs_torch.MinvF_float_fastGrad(Md, Fd, Vs, 3.0, 500, 1e-10, True, True)
torch.cuda.synchronize()
ref = oracle.minvf(M.astype(np.float64), F, sc.y.astype(np.float32).astype(np.float64), 3.0, 500, 1e-10)
res["rel"] = float(np.linalg.norm(X.detach().cpu().numpy() - ref.x) / np.linalg.norm(ref.x))
W = torch.randn(sc.n, 2, device=dev, generator=torch.Generator(dev).manual_seed(1))
gM, gF = torch.autograd.grad((X * W).sum(), (Md, Fd))
M64, F64, G = M.astype(np.float64), F.astype(np.float64), W.cpu().numpy().astype(np.float64)
N = M64.T @ M64
A = np.linalg.solve(N, M64.T @ F64)
V = np.linalg.solve(N, G)
res["gF"] = float(np.linalg.norm(gF.cpu().numpy() - M64 @ V) / np.linalg.norm(M64 @ V))
gM_ref = F64 @ V.T - M64 @ (A @ V.T + V @ A.T)
res["gM"] = float(np.linalg.norm(gM.cpu().numpy() - gM_ref) / np.linalg.norm(gM_ref))
print("RESULT " + json.dumps(res), flush=True)
""")
    lines = out.strip().splitlines()
    got = json.loads([ln for ln in lines if ln.startswith("RESULT ")][-1][7:])
    print(out, flush=True)
    assert "Please run MinvF_float first" in got["vpl_first"]
    assert got["rel"] < 1e-5
    assert got["gF"] < 1e-4 and got["gM"] < 1e-4
    phases = [ln.split("(")[1].split(")")[0] for ln in lines if ln.startswith("TIME(")]
    assert phases[:11] == ["prep", "fps_sampler", "permutation", "build_cov_mat", "pattern", "setup",
                           "compute", "cache", "pcg", "perm_back", "total"]
    assert any(ln.startswith("pcg column 1: iterations") for ln in lines)
