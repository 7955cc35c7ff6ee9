"""Drop-in for fastMBCUI's ``core/mfs_torch_py`` package.

The reference package loads ``./build/mfs_torch/csrc/libmfs_torch.so`` to get
the ``mfs_torch`` operators (core/mfs_torch_py/__init__.py:7-8).  Here the
operators are registered by ``ops`` itself and run on ``libmfs_core.so``
through its C-ABI; importing fails loudly if that library was not built.
"""
from .. import _lib as _core

_core.lib()  # no CPU fallback: fail at import if the HIP library is absent

from . import ops  # noqa: E402
from .ops import MinvF_float_fastGrad, VpL_fast_float, last_solve_info  # noqa: E402,F401
