"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every
symbol include/*.h declares, and the operator surface has the
reference's names, schemas and fake kernels (mfs_torch/csrc/mfs_torch_ops.cpp:7-10,
core/mfs_torch_py/ops.py:12-21, 63-72).  No device calls."""
import ctypes
import os
import re

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    syms = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms.update(re.findall(r"\b(mfs_[a-z0-9_]+)\s*\(", text))
    return sorted(syms)


def test_library_exports_every_declared_symbol():
    from fastmbcui_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from fastmbcui_amd import _lib
    L = _lib.lib()
    for s in _declared_symbols():
        assert getattr(L, s).argtypes is not None or getattr(L, s).restype is None, s


def test_operator_schemas_match_reference():
    import fastmbcui_amd.mfs_torch_py  # noqa: F401  (registers the ops)
    want = {
        "VpL_fast_float": "mfs_torch::VpL_fast_float(Tensor M, Tensor pL, Tensor Vs_s, float rho, "
                          "int pcg_max_iter, float tol_r, bool timing_verbose=False, "
                          "bool solve_verbose=False) -> Tensor",
        "MinvF_float_fastGrad": "mfs_torch::MinvF_float_fastGrad(Tensor M, Tensor F, Tensor Vs_s, "
                                "float rho, int pcg_max_iter, float tol_r, bool timing_verbose=False, "
                                "bool solve_verbose=False) -> Tensor",
    }
    for name, schema in want.items():
        assert str(getattr(torch.ops.mfs_torch, name).default._schema) == schema


def test_fake_kernels_give_reference_shapes():
    import fastmbcui_amd.mfs_torch_py as mfs
    from torch._subclasses.fake_tensor import FakeTensorMode
    with FakeTensorMode():
        M = torch.empty(10, 7)
        F = torch.empty(10, 3)
        V = torch.empty(7, 3)
        out = mfs.ops.MinvF_float_fastGrad(M, F, V, 3.0, 10, 1e-6)
        assert tuple(out.shape) == (7, 3)
        g = torch.empty(7, 2)
        out2 = mfs.ops.VpL_fast_float(M, g, V, 3.0, 10, 1e-6)
        assert tuple(out2.shape) == (7, 2)


def test_stale_surface_is_importable_and_aliased():
    import fastmbcui_amd.mfs_torch as old
    assert callable(old.MinvF) and callable(old.Minvb) and callable(old.VpL)
    a = torch.ones(3)
    assert torch.equal(old.test_add(a, a), 2 * a)


def test_cpu_tensors_are_rejected_by_the_cuda_only_op():
    import fastmbcui_amd.mfs_torch_py as mfs
    M = torch.zeros(4, 4)
    F = torch.zeros(4, 1)
    V = torch.zeros(4, 3)
    with pytest.raises(NotImplementedError):
        mfs.ops.MinvF_float_fastGrad(M, F, V, 3.0, 10, 1e-6)
