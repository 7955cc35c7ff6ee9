"""ctypes binding of libmfs_core.so (the C-ABI in include/mfs_core.h).

The product path has no CPU fallback: if the HIP library is missing this module
raises on first use, naming the build command.  Tensors cross the boundary as
raw device pointers plus sizes; streams are torch's current HIP stream.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmfs_core.so")

MFS_OK, MFS_ERR_ARG, MFS_ERR_HIP, MFS_ERR_DUPLICATE, MFS_ERR_NOT_SPD, MFS_ERR_STATE, MFS_ERR_ALLOC = range(7)
MFS_FP64, MFS_FP32 = 0, 1
MFS_RHS_BOUNDARY, MFS_RHS_SOURCE = 0, 1
(MFS_K_NONE, MFS_K_MATVEC_ROWS, MFS_K_MATVEC_COLS, MFS_K_GRAM, MFS_K_RESID_ROWS, MFS_K_RESID_COLS,
 MFS_K_MATVEC_FUSED) = range(7)
(MFS_PARAM_POINTS_BLOCKS, MFS_PARAM_COLS_BLOCKS, MFS_PARAM_FUSED, MFS_PARAM_TILE_SET,
 MFS_PARAM_FACTOR_PHASES, MFS_PARAM_FACTOR_STREAMS, MFS_PARAM_FACTOR_SPLITS,
 MFS_PARAM_GRAM_SPLIT, MFS_PARAM_FACTOR_BANDS) = 1, 2, 3, 4, 5, 6, 7, 8, 9
TILE_SIZES = (16, 12, 8, 6, 4, 3, 2, 1)   # bit i of MFS_PARAM_TILE_SET

_STATUS_NAMES = {1: "MFS_ERR_ARG", 2: "MFS_ERR_HIP", 3: "MFS_ERR_DUPLICATE", 4: "MFS_ERR_NOT_SPD",
                 5: "MFS_ERR_STATE", 6: "MFS_ERR_ALLOC"}


class MfsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("n_super", ctypes.c_int32), ("max_local", ctypes.c_int32),
        ("band_supernodes", ctypes.c_int32),
        ("nnz_pattern", ctypes.c_int64), ("nnz_factor", ctypes.c_int64),
        ("theta_entries", ctypes.c_int64),
        ("sum_ns2", ctypes.c_double), ("sum_ns3", ctypes.c_double),
        ("plan_ms", ctypes.c_double), ("factor_ms", ctypes.c_double), ("solve_ms", ctypes.c_double),
        ("iters", ctypes.c_int32 * 64), ("relres", ctypes.c_double * 64),
        ("outer_iters", ctypes.c_int32), ("retried_supernodes", ctypes.c_int32),
        ("plan_order_ms", ctypes.c_double), ("plan_pattern_ms", ctypes.c_double),
        ("plan_grid_columns", ctypes.c_int32), ("fp64_sweeps", ctypes.c_int32),
        ("reverted_columns", ctypes.c_int32), ("fused_stale_polls", ctypes.c_int32),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("plan_ms", ctypes.c_double), ("factor_ms", ctypes.c_double), ("solve_ms", ctypes.c_double),
        ("factor_launches", ctypes.c_int32), ("factor_items", ctypes.c_int32),
        ("graph_nodes", ctypes.c_int32), ("graph_replays", ctypes.c_int32),
        ("sweeps", ctypes.c_int32), ("graph_captures", ctypes.c_int32),
    ]


_lib = None
_lock = threading.Lock()


def lib():
    """Load libmfs_core.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m fastmbcui_amd.build` "
                              "(hipcc, gfx950); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        ci = ctypes.c_int
        sig = {
            "mfs_ctx_create": (ci, [ctypes.POINTER(vp), ci, vp]),
            "mfs_ctx_destroy": (None, [vp]),
            "mfs_ctx_set_stream": (ci, [vp, vp]),
            "mfs_ctx_last_error": (ctypes.c_char_p, [vp]),
            "mfs_ctx_stats": (ci, [vp, ctypes.POINTER(Stats)]),
            "mfs_plan": (ci, [vp, vp, i64, f64, f64, i32]),
            "mfs_plan_f64": (ci, [vp, vp, i64, f64, f64, i32]),
            "mfs_plan_export": (ci, [vp] + [vp] * 10),
            "mfs_factor_dense": (ci, [vp, vp, i64, i64, ci]),
            "mfs_factor_export": (ci, [vp, vp]),
            "mfs_solve_dense": (ci, [vp, vp, i64, i64, vp, i64, ci, vp, i32, f64]),
            "mfs_solve_dense64": (ci, [vp, vp, i64, i64, vp, i64, ci, vp, i32, f64]),
            "mfs_order_maximin": (ci, [vp, i64, vp, vp, vp]),
            "mfs_solve_mixed": (ci, [vp, vp, i64, i64, vp, i64, ci, vp, vp, i32, f64, f64]),
            "mfs_normal_residual": (ci, [vp, vp, i64, i64, vp, i64, vp, vp]),
            "mfs_assemble_dense": (ci, [vp, vp, vp, vp, i64, i64, vp]),
            "mfs_factor_points": (ci, [vp, vp, vp, i64, vp, i64, ci]),
            "mfs_solve_points": (ci, [vp, vp, i64, vp, vp, i32, f64, f64]),
            "mfs_points_residual": (ci, [vp, vp, i64, vp, vp]),
            "mfs_solve_points64": (ci, [vp, vp, i64, vp, vp, i32, f64, f64]),
            "mfs_points_residual64": (ci, [vp, vp, i64, vp, vp]),
            "mfs_points_apply": (ci, [vp, ci, ci, vp, i64, vp]),
            "mfs_dense_normal_apply": (ci, [vp, vp, i64, i64, vp, i64, vp]),
            "mfs_dense_normal_apply_noalloc": (ci, [vp, vp, i64, i64, vp, i64, vp]),
            "mfs_dense_grad": (ci, [vp, vp, i64, i64, vp, vp, vp, i64, vp, vp]),
            "mfs_solid_angle": (ci, [vp, ci, vp, i64, vp, i64, vp, i64, vp]),
            "mfs_solid_angle_backward": (ci, [vp, ci, vp, i64, vp, i64, vp, i64, vp, vp, vp]),
            "mfs_single_layer_apply": (ci, [vp, ci, vp, i64, vp, i64, vp, i64, vp]),
            "mfs_single_layer_backward": (ci, [vp, ci, vp, i64, vp, i64, vp, i64, vp, vp, vp, vp]),
            "mfs_batch_create": (ci, [ctypes.POINTER(vp), ci, vp, i32]),
            "mfs_batch_destroy": (None, [vp]),
            "mfs_batch_scene": (vp, [vp, i32]),
            "mfs_batch_last_error": (ctypes.c_char_p, [vp]),
            "mfs_batch_stats_get": (ci, [vp, ctypes.POINTER(BatchStats)]),
            "mfs_batch_plan": (ci, [vp, vp, vp, f64, f64, i32, i32]),
            "mfs_batch_factor_points": (ci, [vp, vp, vp, vp, vp, ci]),
            "mfs_batch_set_factor_mode": (ci, [vp, i32]),
            "mfs_batch_solve_points": (ci, [vp, vp, vp, vp, i32, f64, f64, i32, ci]),
            "mfs_ctx_set_param": (ci, [vp, ci, i64]),
            "mfs_mlp_create": (ci, [ctypes.POINTER(vp), ci, vp, ctypes.c_uint64]),
            "mfs_mlp_destroy": (None, [vp]),
            "mfs_mlp_last_error": (ctypes.c_char_p, [vp]),
            "mfs_mlp_set_stream": (ci, [vp, vp]),
            "mfs_mlp_param_count": (i64, [vp]),
            "mfs_mlp_feature_dim": (i32, []),
            "mfs_mlp_get_params": (ci, [vp, vp]),
            "mfs_mlp_set_params": (ci, [vp, vp]),
            "mfs_mlp_get_grads": (ci, [vp, vp]),
            "mfs_mlp_features": (ci, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i64, vp, vp]),
            "mfs_mlp_loss_grad": (ci, [vp, vp, vp, vp, i64, ctypes.c_float, ctypes.POINTER(f64)]),
            "mfs_mlp_adam_step": (ci, [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
            "mfs_mlp_train_step": (ci, [vp, vp, vp, vp, i64, ctypes.c_float, ctypes.c_float,
                                        ctypes.POINTER(f64)]),
            "mfs_mlp_predict": (ci, [vp, vp, vp, i64, vp]),
            "mfs_mlp_fit_step": (ci, [vp, vp, vp, i32, vp, vp, vp, vp, vp, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ci, ctypes.POINTER(f64)]),
            "mfs_mlp_infer": (ci, [vp, vp, vp, i64, vp]),
            "mfs_ctx_track_kernel": (ci, [vp, ci]),
            "mfs_ctx_tracked": (ci, [vp, ci, ctypes.POINTER(i64), ctypes.POINTER(f64), ctypes.POINTER(f64)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(t):
    """Device (or host) address of a tensor / ndarray, or None."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(t.data_ptr())


def _stream_handle(device):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class Context:
    """One plan + one factor on one device (mfs_ctx)."""

    def __init__(self, device=0):
        import torch
        self.device = torch.device("cuda", device if isinstance(device, int) else device.index or 0)
        h = ctypes.c_void_p()
        self._check(lib().mfs_ctx_create(ctypes.byref(h), self.device.index,
                                         _stream_handle(self.device)), "mfs_ctx_create", raw=True)
        self._h = h
        self.n = None
        self.m_points = None
        self.factor_precision = None

    def close(self):
        if getattr(self, "_h", None):
            lib().mfs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code, what, raw=False):
        if code != MFS_OK:
            msg = what if raw else lib().mfs_ctx_last_error(self._h).decode()
            raise MfsError(code, msg)

    def _sync_stream(self):
        lib().mfs_ctx_set_stream(self._h, _stream_handle(self.device))

    # ---- plan ---------------------------------------------------------------
    def plan(self, xyz, rho, lam=1.5, max_super=2048):
        """xyz: host n x 3 (float32 numpy/tensor -> mfs_plan, float64 -> mfs_plan_f64)."""
        import torch
        self._sync_stream()
        if isinstance(xyz, torch.Tensor):
            xyz = xyz.detach().cpu().numpy()
        xyz = np.ascontiguousarray(xyz)
        if xyz.ndim != 2 or xyz.shape[1] != 3:
            raise ValueError("source points must be n x 3")
        if xyz.dtype == np.float32:
            code = lib().mfs_plan(self._h, _ptr(xyz), xyz.shape[0], float(rho), float(lam), int(max_super))
        else:
            xyz = xyz.astype(np.float64, copy=False)
            code = lib().mfs_plan_f64(self._h, _ptr(xyz), xyz.shape[0], float(rho), float(lam), int(max_super))
        self._check(code, "mfs_plan")
        self.n = int(xyz.shape[0])
        self.factor_precision = None

    def plan_export(self):
        s = self.stats()
        n, ns = s.n, s.n_super
        out = dict(P=np.empty(n, np.int32), Pinv=np.empty(n, np.int32), scale=np.empty(n, np.float64),
                   ptr=np.empty(n + 1, np.int32), ind=np.empty(max(s.nnz_pattern, 1), np.int32),
                   sup_ptr=np.empty(ns + 1, np.int32), sup_ind=np.empty(n, np.int32),
                   sup_parent=np.empty(n, np.int32), sptr=np.empty(n + 1, np.int32),
                   sind=np.empty(max(s.nnz_factor, 1), np.int32))
        keys = ["P", "Pinv", "scale", "ptr", "ind", "sup_ptr", "sup_ind", "sup_parent", "sptr", "sind"]
        self._check(lib().mfs_plan_export(self._h, *[_ptr(out[k]) for k in keys]), "mfs_plan_export")
        out["ind"] = out["ind"][: s.nnz_pattern]
        out["sind"] = out["sind"][: s.nnz_factor]
        return out

    # ---- factor -------------------------------------------------------------
    def factor_dense(self, M, precision=MFS_FP64):
        self._sync_stream()
        _check_dense(M, self.n)
        self._check(lib().mfs_factor_dense(self._h, _ptr(M), M.shape[0], M.shape[1], int(precision)),
                    "mfs_factor_dense")
        self.factor_precision = precision

    def factor_export(self):
        s = self.stats()
        val = np.empty(max(s.nnz_factor, 1), np.float64)
        self._check(lib().mfs_factor_export(self._h, _ptr(val)), "mfs_factor_export")
        return val[: s.nnz_factor]

    # ---- solves -------------------------------------------------------------
    def solve_dense(self, M, B, rhs_space=MFS_RHS_BOUNDARY, maxit=100, tol=1e-6, out=None):
        """Plain FP64 PCG per right-hand side.  B: m x k (or n x k), float32 cuda."""
        import torch
        self._sync_stream()
        _check_dense(M, self.n)
        B = _as_2d_f32(B)
        if out is None:
            out = torch.empty((M.shape[1], B.shape[1]), dtype=torch.float32, device=M.device)
        self._check(lib().mfs_solve_dense(self._h, _ptr(M), M.shape[0], M.shape[1], _ptr(B), B.shape[1],
                                          int(rhs_space), _ptr(out), int(maxit), float(tol)),
                    "mfs_solve_dense")
        return out

    def solve_dense64(self, M, B, rhs_space=MFS_RHS_BOUNDARY, maxit=100, tol=1e-6, out=None):
        """solve_dense with FP64 strengths (n x k float64 cuda)."""
        import torch
        self._sync_stream()
        _check_dense(M, self.n)
        B = _as_2d_f32(B)
        if out is None:
            out = torch.empty((M.shape[1], B.shape[1]), dtype=torch.float64, device=M.device)
        self._check(lib().mfs_solve_dense64(self._h, _ptr(M), M.shape[0], M.shape[1], _ptr(B),
                                            B.shape[1], int(rhs_space), _ptr(out), int(maxit),
                                            float(tol)), "mfs_solve_dense64")
        return out

    def solve_mixed(self, M, B, maxit=1000, tol=1e-6, inner_tol=0.5, x0=None, out=None,
                    rhs_space=None):
        """Batched FP32 PCG with FP64 refinement.  B: m x k float32 cuda
        (MFS_RHS_BOUNDARY, the default) or n x k (MFS_RHS_SOURCE, VpL)."""
        if rhs_space is None:
            rhs_space = MFS_RHS_BOUNDARY
        import torch
        self._sync_stream()
        _check_dense(M, self.n)
        B = _as_2d_f32(B)
        if out is None:
            out = torch.empty((M.shape[1], B.shape[1]), dtype=torch.float32, device=M.device)
        if x0 is not None:
            x0 = _as_2d_f32(x0)
        rows = M.shape[0] if rhs_space == MFS_RHS_BOUNDARY else M.shape[1]
        if B.shape[0] != rows:
            raise ValueError("right-hand sides have the wrong number of rows")
        self._check(lib().mfs_solve_mixed(self._h, _ptr(M), M.shape[0], M.shape[1], _ptr(B), B.shape[1],
                                          int(rhs_space), _ptr(x0), _ptr(out), int(maxit), float(tol),
                                          float(inner_tol)), "mfs_solve_mixed")
        return out

    def normal_residual(self, M, B, X):
        self._sync_stream()
        B = _as_2d_f32(B)
        X = _as_2d_f32(X)
        out = np.empty(B.shape[1], np.float64)
        self._check(lib().mfs_normal_residual(self._h, _ptr(M), M.shape[0], M.shape[1], _ptr(B),
                                              B.shape[1], _ptr(X), _ptr(out)), "mfs_normal_residual")
        return out

    def assemble_dense(self, x, nx, y):
        """M (float32, m x n) from FP64 device points x, nx (m x 3) and y (n x 3)."""
        import torch
        self._sync_stream()
        M = torch.empty((x.shape[0], y.shape[0]), dtype=torch.float32, device=x.device)
        self._check(lib().mfs_assemble_dense(self._h, _ptr(x.contiguous()), _ptr(nx.contiguous()),
                                             _ptr(y.contiguous()), x.shape[0], y.shape[0], _ptr(M)),
                    "mfs_assemble_dense")
        return M

    # ---- points (matrix-free) form ---------------------------------------------
    def factor_points(self, x, nx, y, precision=MFS_FP32):
        """Pack FP64 device points (x, nx: m x 3 targets/normals; y: n x 3 planned
        sources, original order) and factor with the Gram evaluated from them."""
        import torch
        self._sync_stream()
        pts = [t.to(torch.float64).contiguous() for t in (x, nx, y)]
        for t in pts:
            if not t.is_cuda or t.dim() != 2 or t.shape[1] != 3:
                raise ValueError("points must be (count, 3) device tensors")
        if pts[0].shape != pts[1].shape or self.n is None or pts[2].shape[0] != self.n:
            raise ValueError("x and nx must match; y must hold the planned sources")
        self._points = pts  # keep the inputs alive until the packing has run
        self.m_points = pts[0].shape[0]
        self._check(lib().mfs_factor_points(self._h, _ptr(pts[0]), _ptr(pts[1]), pts[0].shape[0],
                                            _ptr(pts[2]), pts[2].shape[0], int(precision)),
                    "mfs_factor_points")
        self.factor_precision = precision

    def solve_points(self, B, maxit=1000, tol=1e-6, inner_tol=0.5, x0=None, out=None):
        """Matrix-free mixed solve.  B: m x k float32 cuda."""
        import torch
        self._sync_stream()
        B = _as_2d_f32(B)
        if B.shape[0] != self.m_points:
            raise ValueError("B must have one row per target point")
        if out is None:
            out = torch.empty((self.n, B.shape[1]), dtype=torch.float32, device=B.device)
        if x0 is not None:
            x0 = _as_2d_f32(x0)
        self._check(lib().mfs_solve_points(self._h, _ptr(B), B.shape[1], _ptr(x0), _ptr(out), int(maxit),
                                           float(tol), float(inner_tol)), "mfs_solve_points")
        return out

    def solve_points64(self, B, maxit=1000, tol=1e-6, inner_tol=0.5, x0=None, out=None):
        """solve_points with FP64 warm start and FP64 strengths (n x k float64 cuda)."""
        import torch
        self._sync_stream()
        B = _as_2d_f32(B)
        if B.shape[0] != self.m_points:
            raise ValueError("B must have one row per target point")
        if out is None:
            out = torch.empty((self.n, B.shape[1]), dtype=torch.float64, device=B.device)
        if x0 is not None:
            x0 = (x0.unsqueeze(1) if x0.dim() == 1 else x0).to(torch.float64).contiguous()
        self._check(lib().mfs_solve_points64(self._h, _ptr(B), B.shape[1], _ptr(x0), _ptr(out), int(maxit),
                                             float(tol), float(inner_tol)), "mfs_solve_points64")
        return out

    def points_residual(self, B, X):
        import torch
        self._sync_stream()
        B = _as_2d_f32(B)
        X = X.unsqueeze(1) if X.dim() == 1 else X
        out = np.empty(B.shape[1], np.float64)
        if X.dtype == torch.float64:
            X = X.contiguous()
            self._check(lib().mfs_points_residual64(self._h, _ptr(B), B.shape[1], _ptr(X), _ptr(out)),
                        "mfs_points_residual64")
        else:
            X = _as_2d_f32(X)
            self._check(lib().mfs_points_residual(self._h, _ptr(B), B.shape[1], _ptr(X), _ptr(out)),
                        "mfs_points_residual")
        return out

    def dense_normal_apply(self, M, X):
        """Q = M^T (M X) in FP64 (X: n x k float32 cuda, k <= 16) through the
        PCG's matvec (one pass over M when MFS_PARAM_FUSED allows it)."""
        import torch
        self._sync_stream()
        if X.dim() == 1:
            X = X.unsqueeze(1)
        X = X.to(torch.float32).contiguous()
        if M.dim() != 2 or X.shape[0] != M.shape[1]:
            raise ValueError("X must have as many rows as M has columns")
        Q = torch.empty((M.shape[1], X.shape[1]), dtype=torch.float64, device=X.device)
        self._check(lib().mfs_dense_normal_apply(self._h, _ptr(M.contiguous()), M.shape[0], M.shape[1],
                                                 _ptr(X), X.shape[1], _ptr(Q)), "mfs_dense_normal_apply")
        return Q

    def dense_normal_apply_noalloc(self, M, X, Q=None):
        """dense_normal_apply through the entry point that allocates nothing
        (static device scratch); asynchronous on the current stream.  X must be
        a contiguous n x k float32 device tensor, Q (n x k float64) is
        allocated here when not given."""
        import torch
        self._sync_stream()
        if X.dim() == 1:
            X = X.unsqueeze(1)
        if X.dtype != torch.float32 or not X.is_contiguous() or not M.is_contiguous():
            raise ValueError("M and X must be contiguous float32 device tensors")
        if M.dim() != 2 or X.shape[0] != M.shape[1]:
            raise ValueError("X must have as many rows as M has columns")
        if Q is None:
            Q = torch.empty((M.shape[1], X.shape[1]), dtype=torch.float64, device=X.device)
        self._check(lib().mfs_dense_normal_apply_noalloc(self._h, _ptr(M), M.shape[0], M.shape[1],
                                                         _ptr(X), X.shape[1], _ptr(Q)),
                    "mfs_dense_normal_apply_noalloc")
        return Q

    def dense_grad(self, M, F, A, V, need_m=True, need_f=True):
        """(grad_M, grad_F) of the solve's backward: grad_F = M V,
        grad_M = (F - M A) V^T - (M V) A^T; FP32 device tensors, k <= 16."""
        import torch
        self._sync_stream()
        F, A, V = (_as_2d_f32(t) for t in (F, A, V))
        if not M.is_contiguous() or M.dtype != torch.float32:
            raise ValueError("M must be a contiguous float32 device tensor")
        m, n = M.shape
        k = F.shape[1]
        if F.shape[0] != m or A.shape != (n, k) or V.shape != (n, k):
            raise ValueError("F must be m x k, A and V n x k")
        gM = torch.empty((m, n), dtype=torch.float32, device=M.device) if need_m else None
        gF = torch.empty((m, k), dtype=torch.float32, device=M.device) if need_f else None
        self._check(lib().mfs_dense_grad(self._h, _ptr(M), m, n, _ptr(F), _ptr(A), _ptr(V), k,
                                         _ptr(gM), _ptr(gF)), "mfs_dense_grad")
        return gM, gF

    def points_apply(self, v, trans=False, precision=MFS_FP32):
        """M v (v: n x k) or M^T v (trans, v: m x k) from the packed points."""
        import torch
        self._sync_stream()
        dt = torch.float64 if precision == MFS_FP64 else torch.float32
        if v.dim() == 1:
            v = v.unsqueeze(1)
        v = v.to(dt).contiguous()
        rows_in, rows_out = (self.m_points, self.n) if trans else (self.n, self.m_points)
        if v.shape[0] != rows_in:
            raise ValueError("operand has the wrong number of rows")
        out = torch.empty((rows_out, v.shape[1]), dtype=dt, device=v.device)
        self._check(lib().mfs_points_apply(self._h, int(bool(trans)), int(precision), _ptr(v), v.shape[1],
                                           _ptr(out)), "mfs_points_apply")
        return out

    # ---- caller steps: face solid angles and the single-layer potential ------
    @staticmethod
    def _caller_precision(t):
        import torch
        if t.dtype == torch.float32:
            return MFS_FP32
        if t.dtype == torch.float64:
            return MFS_FP64
        raise TypeError("expected a float32 or float64 tensor")

    def solid_angle(self, V, F, S):
        """M (nf x n): solid angle of each face of (V, F) seen from each source S."""
        import torch
        self._sync_stream()
        prec = self._caller_precision(V)
        V, S = V.contiguous(), S.to(V.dtype).contiguous()
        F = F.to(torch.int32).contiguous()
        _check_xyz(V, "V"), _check_xyz(S, "S"), _check_faces(F)
        M = torch.empty((F.shape[0], S.shape[0]), dtype=V.dtype, device=V.device)
        self._check(lib().mfs_solid_angle(self._h, prec, _ptr(V), V.shape[0], _ptr(F), F.shape[0],
                                          _ptr(S), S.shape[0], _ptr(M)), "mfs_solid_angle")
        return M

    def solid_angle_backward(self, V, F, S, gM, need_v=True, need_s=True):
        """(dL/dV, dL/dS) from dL/dM; either is None when not asked for."""
        import torch
        self._sync_stream()
        prec = self._caller_precision(V)
        V, S = V.contiguous(), S.to(V.dtype).contiguous()
        F = F.to(torch.int32).contiguous()
        gM = gM.to(V.dtype).contiguous()
        _check_xyz(V, "V"), _check_xyz(S, "S"), _check_faces(F)
        if tuple(gM.shape) != (F.shape[0], S.shape[0]):
            raise ValueError("gM must be nf x n")
        gV = torch.empty_like(V) if need_v else None
        gS = torch.empty_like(S) if need_s else None
        self._check(lib().mfs_solid_angle_backward(self._h, prec, _ptr(V), V.shape[0], _ptr(F),
                                                   F.shape[0], _ptr(S), S.shape[0], _ptr(gM),
                                                   _ptr(gV), _ptr(gS)), "mfs_solid_angle_backward")
        return gV, gS

    def single_layer_apply(self, S, C, X):
        """phi (nf x k) = SL X, SL[f][s] = 1 / |C[f] - S[s]|, without storing SL."""
        import torch
        self._sync_stream()
        prec = self._caller_precision(X)
        S, C, X = S.to(X.dtype).contiguous(), C.to(X.dtype).contiguous(), X.contiguous()
        _check_xyz(S, "S"), _check_xyz(C, "C")
        if X.dim() != 2 or X.shape[0] != S.shape[0]:
            raise ValueError("X must be n x k")
        phi = torch.empty((C.shape[0], X.shape[1]), dtype=X.dtype, device=X.device)
        self._check(lib().mfs_single_layer_apply(self._h, prec, _ptr(S), S.shape[0], _ptr(C),
                                                 C.shape[0], _ptr(X), X.shape[1], _ptr(phi)),
                    "mfs_single_layer_apply")
        return phi

    def single_layer_backward(self, S, C, X, gphi, need_s=True, need_c=True, need_x=True):
        """(dL/dS, dL/dC, dL/dX) from dL/dphi; entries not asked for are None."""
        import torch
        self._sync_stream()
        prec = self._caller_precision(X)
        S, C, X = S.to(X.dtype).contiguous(), C.to(X.dtype).contiguous(), X.contiguous()
        gphi = gphi.to(X.dtype).contiguous()
        if tuple(gphi.shape) != (C.shape[0], X.shape[1]):
            raise ValueError("gphi must be nf x k")
        gS = torch.empty_like(S) if need_s else None
        gC = torch.empty_like(C) if need_c else None
        gX = torch.empty_like(X) if need_x else None
        self._check(lib().mfs_single_layer_backward(self._h, prec, _ptr(S), S.shape[0], _ptr(C),
                                                    C.shape[0], _ptr(X), X.shape[1], _ptr(gphi),
                                                    _ptr(gS), _ptr(gC), _ptr(gX)),
                    "mfs_single_layer_backward")
        return gS, gC, gX

    def set_param(self, param, value):
        """Launch-shape parameter (MFS_PARAM_*; 0 = default)."""
        self._check(lib().mfs_ctx_set_param(self._h, int(param), int(value)), "mfs_ctx_set_param")

    # ---- statistics -----------------------------------------------------------
    def stats(self):
        s = Stats()
        self._check(lib().mfs_ctx_stats(self._h, ctypes.byref(s)), "mfs_ctx_stats")
        return s

    def track(self, classes):
        """Enable per-launch event timing for the given kernel classes (iterable of ints)."""
        mask = 0
        for c in classes:
            mask |= 1 << int(c)
        self._check(lib().mfs_ctx_track_kernel(self._h, mask), "mfs_ctx_track_kernel")

    def tracked(self, kernel_class):
        """(launches, summed event ms, algorithmic bytes) for one kernel class."""
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0)
        by = ctypes.c_double(0)
        self._check(lib().mfs_ctx_tracked(self._h, int(kernel_class), ctypes.byref(n), ctypes.byref(ms),
                                          ctypes.byref(by)), "mfs_ctx_tracked")
        return n.value, ms.value, by.value


def order_maximin(xyz):
    """Host maximin ordering of n x 3 points: (P, Pinv, scale), as the plan
    computes it (no device work)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    if xyz.ndim != 2 or xyz.shape[1] != 3:
        raise ValueError("points must be n x 3")
    n = xyz.shape[0]
    P = np.empty(n, np.int32)
    Pinv = np.empty(n, np.int32)
    scale = np.empty(n, np.float64)
    code = lib().mfs_order_maximin(_ptr(xyz), n, _ptr(P), _ptr(Pinv), _ptr(scale))
    if code != MFS_OK:
        raise MfsError(code, "mfs_order_maximin")
    return P, Pinv, scale


def _check_xyz(t, name):
    if t.dim() != 2 or t.shape[1] != 3:
        raise ValueError(f"{name} must be (count, 3)")


def _check_faces(F):
    if F.dim() != 2 or F.shape[1] != 3:
        raise ValueError("F must be (num_faces, 3)")


def _check_dense(M, n):
    import torch
    if not isinstance(M, torch.Tensor) or M.dtype != torch.float32 or not M.is_cuda or M.dim() != 2:
        raise ValueError("M must be a 2-D float32 device tensor")
    if not M.is_contiguous():
        raise ValueError("M must be contiguous (row-major)")
    if n is None or M.shape[1] != n:
        raise ValueError("M's column count must equal the planned source count")


def _as_2d_f32(B):
    import torch
    if B.dim() == 1:
        B = B.unsqueeze(1)
    return B.to(torch.float32).contiguous()


def _ptr_array(items):
    arr = (ctypes.c_void_p * len(items))()
    for i, t in enumerate(items):
        arr[i] = None if t is None else (t.ctypes.data if isinstance(t, np.ndarray) else t.data_ptr())
    return arr


class Batch:
    """Independent scenes solved together (mfs_batch): parallel host plans, one
    merged factor launch per tile size, PCG iterations launched in rounds on
    the scenes' streams."""

    def __init__(self, count, device=0):
        import torch
        self.device = torch.device("cuda", device)
        self.count = int(count)
        h = ctypes.c_void_p()
        code = lib().mfs_batch_create(ctypes.byref(h), self.device.index, _stream_handle(self.device),
                                      self.count)
        if code != MFS_OK:
            raise MfsError(code, "mfs_batch_create")
        self._h = h
        self.n = [None] * self.count
        self.m = [None] * self.count
        self._keep = None

    def close(self):
        if getattr(self, "_h", None):
            lib().mfs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code, what):
        if code != MFS_OK:
            raise MfsError(code, f"{what}: {lib().mfs_batch_last_error(self._h).decode()}")

    @staticmethod
    def _sync():
        import torch
        torch.cuda.synchronize()

    def plan(self, sources, rho, lam=1.5, max_super=2048, threads=8):
        """sources: list of (n_i, 3) host arrays (FP64)."""
        pts = [np.ascontiguousarray(p, dtype=np.float64) for p in sources]
        n = np.array([p.shape[0] for p in pts], dtype=np.int64)
        self._check(lib().mfs_batch_plan(self._h, _ptr_array(pts), n.ctypes.data, float(rho), float(lam),
                                         int(max_super), int(threads)), "mfs_batch_plan")
        self.n = [int(v) for v in n]

    def set_factor_mode(self, mode):
        """2: every scene's supernodes in one launch over a device queue;
        1: one launch per tile size and row split count; 0 (default): the
        queue launch for up to 32,768 units, else a launch per bucket."""
        self._check(lib().mfs_batch_set_factor_mode(self._h, int(mode)), "mfs_batch_set_factor_mode")

    def factor_points(self, xs, nxs, ys, precision=MFS_FP32):
        """Device FP64 (m_i, 3) targets/normals and (n_i, 3) planned sources per scene."""
        import torch
        xs = [t.to(torch.float64).contiguous() for t in xs]
        nxs = [t.to(torch.float64).contiguous() for t in nxs]
        ys = [t.to(torch.float64).contiguous() for t in ys]
        m = np.array([t.shape[0] for t in xs], dtype=np.int64)
        self._keep = (xs, nxs, ys)
        self._sync()
        self._check(lib().mfs_batch_factor_points(self._h, _ptr_array(xs), _ptr_array(nxs), m.ctypes.data,
                                                  _ptr_array(ys), int(precision)), "mfs_batch_factor_points")
        self.m = [int(v) for v in m]

    def solve_points(self, Bs, maxit=1000, tol=1e-6, inner_tol=0.5, chunk=4, use_graph=0):
        """Bs: list of (m_i, k_i) float32 device tensors, k_i <= 16.  Returns the X_i.
        use_graph: 0 direct launches, 2 per-scene graphs of `chunk` iterations;
        1 (True; one multi-stream graph) is switched off in the library: MfsError."""
        import torch
        Bs = [_as_2d_f32(b) for b in Bs]
        Xs = [torch.empty((self.n[i], b.shape[1]), dtype=torch.float32, device=b.device)
              for i, b in enumerate(Bs)]
        k = np.array([b.shape[1] for b in Bs], dtype=np.int64)
        self._sync()
        self._check(lib().mfs_batch_solve_points(self._h, _ptr_array(Bs), k.ctypes.data, _ptr_array(Xs),
                                                 int(maxit), float(tol), float(inner_tol), int(chunk),
                                                 int(use_graph)), "mfs_batch_solve_points")
        return Xs

    def stats(self):
        s = BatchStats()
        self._check(lib().mfs_batch_stats_get(self._h, ctypes.byref(s)), "mfs_batch_stats_get")
        return s

    def set_param(self, param, value):
        """mfs_ctx_set_param on every scene's context."""
        for i in range(self.count):
            code = lib().mfs_ctx_set_param(lib().mfs_batch_scene(self._h, i), int(param), int(value))
            if code != MFS_OK:
                raise MfsError(code, "mfs_ctx_set_param")

    def scene_stats(self, i):
        s = Stats()
        code = lib().mfs_ctx_stats(lib().mfs_batch_scene(self._h, int(i)), ctypes.byref(s))
        if code != MFS_OK:
            raise MfsError(code, "mfs_ctx_stats")
        return s

    def scene_factor_export(self, i):
        """The factor values of scene i (FP64 host copy, mfs_factor_export)."""
        s = self.scene_stats(i)
        val = np.empty(max(s.nnz_factor, 1), np.float64)
        code = lib().mfs_factor_export(lib().mfs_batch_scene(self._h, int(i)), _ptr(val))
        if code != MFS_OK:
            raise MfsError(code, "mfs_factor_export")
        return val[: s.nnz_factor]
