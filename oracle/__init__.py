"""CPU FP64 oracle for the preconditioned MFS solve -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this package, and only as the checker.  The product
(``fastmbcui_amd``) never imports it.

The C restatement lives in ``mfs_oracle.c`` (built by ``make -C oracle`` into
``oracle/build/libmfs_oracle.so``); this module binds it with ctypes and chains
the steps the way ``mfs_torch/csrc/mfs_torch_ops.cu:21-197`` (MinvF_float) and
``:199-370`` (VpL_fast_float) chain them, in FP64.

What pins each function is listed in ``mfs_oracle.h``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libmfs_oracle.so")

MO_OK, MO_ERR_ARG, MO_ERR_ALLOC, MO_ERR_DUPLICATE, MO_ERR_NOT_SPD = 0, 1, 2, 3, 4

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_i32, c_i64, c_f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        vp = ctypes.c_void_p
        L.mo_fast_fps.argtypes = [c_i32, _f64p, _i32p, _i32p, _f64p, vp, vp]
        L.mo_maximin_brute.argtypes = [c_i32, _f64p, c_i32, _i32p, _f64p]
        L.mo_pattern.argtypes = [c_i32, _f64p, _f64p, c_f64, _i32p, vp]
        L.mo_aggregate.argtypes = [c_i32, _i32p, _i32p, _f64p, c_f64, c_i32, _i32p, _i32p, _i32p,
                                   ctypes.POINTER(c_i32)]
        L.mo_super_pattern.argtypes = [c_i32, _i32p, _i32p, _i32p, c_i32, _i32p, vp]
        L.mo_factor.argtypes = [c_i64, c_i32, _f64p, _i32p, _i32p, c_i32, _i32p, _i32p, _f64p,
                                ctypes.POINTER(c_i32)]
        L.mo_spmv_lt.argtypes = [c_i32, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.mo_spmv_l.argtypes = [c_i32, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.mo_pcg.argtypes = [c_i64, c_i32, _f64p, _f64p, ctypes.c_int, _i32p, _i32p, _f64p,
                             ctypes.c_int, c_i32, c_f64, _f64p, ctypes.POINTER(c_i32),
                             ctypes.POINTER(c_f64)]
        for f in ("mo_fast_fps", "mo_maximin_brute", "mo_pattern", "mo_aggregate",
                  "mo_super_pattern", "mo_factor", "mo_pcg"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed with status {code}")
        self.code = code


def _check(code, what):
    if code != MO_OK:
        raise OracleError(code, what)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------------------
# Plan: ordering, pattern, supernodes
# --------------------------------------------------------------------------------------
def fast_fps(xyz):
    """P (orig->new), Pinv (new->orig), scale (ascending, new order), seq, distances."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    P = np.empty(n, np.int32)
    Pinv = np.empty(n, np.int32)
    scale = np.empty(n, np.float64)
    seq = np.empty(n, np.int32)
    dist = np.empty(n, np.float64)
    _check(lib().mo_fast_fps(n, xyz, P, Pinv, scale, _ptr(seq), _ptr(dist)), "mo_fast_fps")
    return P, Pinv, scale, seq, dist


def maximin_brute(xyz, root):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = xyz.shape[0]
    seq = np.empty(n, np.int32)
    dist = np.empty(n, np.float64)
    _check(lib().mo_maximin_brute(n, xyz, int(root), seq, dist), "mo_maximin_brute")
    return seq, dist


def pattern(xyz_new, scale, rho):
    xyz_new = np.ascontiguousarray(xyz_new, dtype=np.float64)
    scale = np.ascontiguousarray(scale, dtype=np.float64)
    n = xyz_new.shape[0]
    ptr = np.empty(n + 1, np.int32)
    _check(lib().mo_pattern(n, xyz_new, scale, float(rho), ptr, None), "mo_pattern(count)")
    ind = np.empty(max(int(ptr[-1]), 1), np.int32)
    _check(lib().mo_pattern(n, xyz_new, scale, float(rho), ptr, _ptr(ind)), "mo_pattern")
    return ptr, ind[: ptr[-1]]


def aggregate(ptr, ind, scale, lam=1.5, max_super=2048):
    n = ptr.shape[0] - 1
    sup_ptr = np.empty(n + 1, np.int32)
    sup_ind = np.empty(n, np.int32)
    sup_parent = np.empty(n, np.int32)
    ns = ctypes.c_int32(0)
    ind_c = np.ascontiguousarray(ind if ind.size else np.zeros(1, np.int32), dtype=np.int32)
    _check(lib().mo_aggregate(n, np.ascontiguousarray(ptr, np.int32), ind_c,
                              np.ascontiguousarray(scale, np.float64), float(lam), int(max_super),
                              sup_ptr, sup_ind, sup_parent, ctypes.byref(ns)), "mo_aggregate")
    return sup_ptr[: ns.value + 1].copy(), sup_ind, sup_parent


def super_pattern(ptr, ind, sup_parent, n_super):
    n = ptr.shape[0] - 1
    sptr = np.empty(n + 1, np.int32)
    ind_c = np.ascontiguousarray(ind if ind.size else np.zeros(1, np.int32), dtype=np.int32)
    ptr_c = np.ascontiguousarray(ptr, np.int32)
    par = np.ascontiguousarray(sup_parent, np.int32)
    _check(lib().mo_super_pattern(n, ptr_c, ind_c, par, int(n_super), sptr, None),
           "mo_super_pattern(count)")
    sind = np.empty(max(int(sptr[-1]), 1), np.int32)
    _check(lib().mo_super_pattern(n, ptr_c, ind_c, par, int(n_super), sptr, _ptr(sind)),
           "mo_super_pattern")
    return sptr, sind[: sptr[-1]]


@dataclass
class Plan:
    n: int
    rho: float          # as used by the pattern: (double)(float)rho
    P: np.ndarray
    Pinv: np.ndarray
    scale: np.ndarray
    ptr: np.ndarray
    ind: np.ndarray
    sup_ptr: np.ndarray
    sup_ind: np.ndarray
    sup_parent: np.ndarray
    sptr: np.ndarray
    sind: np.ndarray

    @property
    def n_super(self):
        return int(self.sup_ptr.shape[0] - 1)

    def local_sizes(self):
        first = self.sup_ind[self.sup_ptr[:-1]]
        return (self.sptr[first + 1] - self.sptr[first]).astype(np.int64)


def plan(xyz, rho, lam=1.5, max_super=2048):
    """mfs_torch_ops.cu:51-54 and :114-116 with the op's constants (lambda 1.5, cap 2048)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    P, Pinv, scale, _, _ = fast_fps(xyz)
    rho32 = float(np.float32(rho))   # fps.simpl_sparsity((float)rho, ...)
    ptr, ind = pattern(xyz[Pinv], scale, rho32)
    sup_ptr, sup_ind, sup_parent = aggregate(ptr, ind, scale, lam, max_super)
    sptr, sind = super_pattern(ptr, ind, sup_parent, sup_ptr.shape[0] - 1)
    return Plan(xyz.shape[0], rho32, P, Pinv, scale, ptr, ind, sup_ptr, sup_ind, sup_parent,
                sptr, sind)


# --------------------------------------------------------------------------------------
# Factor and PCG
# --------------------------------------------------------------------------------------
def permuted_colmajor(M, Pinv):
    """Mp as an m x n column-major FP64 buffer: column c = M[:, Pinv[c]] (mfs_torch_ops.cu:75,106)."""
    return np.ascontiguousarray(np.asarray(M, dtype=np.float64)[:, Pinv].T)


def factor(Mp_cm, m, pl: Plan):
    val = np.empty(max(int(pl.sptr[-1]), 1), np.float64)
    bad = ctypes.c_int32(-1)
    code = lib().mo_factor(int(m), pl.n, Mp_cm, pl.sup_ptr, pl.sup_ind, pl.n_super, pl.sptr,
                           pl.sind, val, ctypes.byref(bad))
    if code != MO_OK:
        err = OracleError(code, f"mo_factor (supernode {bad.value})")
        err.bad_super = bad.value
        raise err
    return val[: pl.sptr[-1]]


def spmv_lt(pl, val, r):
    t = np.empty(pl.n, np.float64)
    lib().mo_spmv_lt(pl.n, pl.sptr, pl.sind, val, np.ascontiguousarray(r, np.float64), t)
    return t


def spmv_l(pl, val, t):
    z = np.empty(pl.n, np.float64)
    lib().mo_spmv_l(pl.n, pl.sptr, pl.sind, val, np.ascontiguousarray(t, np.float64), z)
    return z


def pcg(Mp_cm, m, pl: Plan, val, rhs, rhs_space=0, maxit=100, tol=1e-6, preconditioned=True):
    """One right-hand side.  Returns (x in new order, iterations, sqrt(RESD2/RHS2))."""
    x = np.empty(pl.n, np.float64)
    it = ctypes.c_int32(0)
    rr = ctypes.c_double(0.0)
    _check(lib().mo_pcg(int(m), pl.n, Mp_cm, np.ascontiguousarray(rhs, np.float64), int(rhs_space),
                        pl.sptr, pl.sind, val, int(bool(preconditioned)), int(maxit), float(tol), x,
                        ctypes.byref(it), ctypes.byref(rr)), "mo_pcg")
    return x, it.value, rr.value


@dataclass
class SolveResult:
    x: np.ndarray          # n x k, original source order
    iters: list
    relres: list
    plan: Plan
    val: np.ndarray


def minvf(M, F, Vs, rho, maxit, tol, pl: Plan | None = None, val=None):
    """FP64 restatement of MinvF_float (mfs_torch_ops.cu:21-197): returns strengths n x k."""
    M = np.asarray(M, dtype=np.float64)
    F = np.asarray(F, dtype=np.float64).reshape(M.shape[0], -1)
    m = M.shape[0]
    pl = pl or plan(Vs, rho)
    Mp = permuted_colmajor(M, pl.Pinv)
    if val is None:
        val = factor(Mp, m, pl)
    X = np.empty((pl.n, F.shape[1]))
    iters, rel = [], []
    for c in range(F.shape[1]):
        xp, it, rr = pcg(Mp, m, pl, val, F[:, c], 0, maxit, tol)
        X[:, c] = xp[pl.P]          # result[i] = result_p[P[i]]  (:178-187)
        iters.append(it)
        rel.append(rr)
    return SolveResult(X, iters, rel, pl, val)


def vpl(M, G, Vs, rho, maxit, tol, pl: Plan | None = None, val=None):
    """FP64 restatement of VpL_fast_float (mfs_torch_ops.cu:199-370): (M^T M)^{-1} G."""
    M = np.asarray(M, dtype=np.float64)
    G = np.asarray(G, dtype=np.float64).reshape(M.shape[1], -1)
    m = M.shape[0]
    pl = pl or plan(Vs, rho)
    Mp = permuted_colmajor(M, pl.Pinv)
    if val is None:
        val = factor(Mp, m, pl)
    X = np.empty((pl.n, G.shape[1]))
    iters, rel = [], []
    for c in range(G.shape[1]):
        xp, it, rr = pcg(Mp, m, pl, val, G[pl.Pinv, c], 1, maxit, tol)
        X[:, c] = xp[pl.P]
        iters.append(it)
        rel.append(rr)
    return SolveResult(X, iters, rel, pl, val)


def normal_residual(M, B, X):
    """max over columns of ||M^T (b - M x)|| / ||M^T b||, FP64 (the measured quantity)."""
    M = np.asarray(M, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64).reshape(M.shape[0], -1)
    X = np.asarray(X, dtype=np.float64).reshape(M.shape[1], -1)
    num = np.linalg.norm(M.T @ (B - M @ X), axis=0)
    den = np.linalg.norm(M.T @ B, axis=0)
    return float(np.max(num / den))
