"""The reference's older ``mfs_torch`` package surface (mfs_torch/__init__.py,
mfs_torch/ops.py), aliased onto the live operators.  The reference's version
declares ``test_add``, ``Minvb``, ``MinvF`` and ``VpL`` but its shared library
never defines them (mfs_torch/csrc/mfs_torch_ops.cpp:7-10), so importing it
fails; here each name maps to the operator it was meant to be."""
from ..mfs_torch_py import ops as _ops  # registers mfs_torch::* (fails loudly without the HIP library)
from . import ops  # noqa: F401
from .ops import Minvb, MinvF, VpL, test_add  # noqa: F401
