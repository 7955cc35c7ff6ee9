"""MI355X-native preconditioned MFS solve (the hot path of fastMBCUI).

* ``fastmbcui_amd.mfs_torch_py`` -- drop-in for fastMBCUI's ``core/mfs_torch_py``
  (``torch.ops.mfs_torch.MinvF_float_fastGrad`` / ``VpL_fast_float``).
* ``fastmbcui_amd.mfs_torch``    -- the reference's older ``mfs_torch`` surface,
  aliased onto the same operators.
* ``fastmbcui_amd.runtime``      -- contexts and the plan/factor cache behind them.
* ``fastmbcui_amd.scenes``       -- seeded synthetic point scenes.
"""
__all__ = ["runtime", "scenes"]
