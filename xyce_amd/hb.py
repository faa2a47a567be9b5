"""Harmonic balance and the batched FP32 LU on the device (torch tensors in,
through the C ABI).

``HBPlan`` holds the sample count N = 2H + 1, the unknowns n of one sample
and the period; its transforms map time-domain samples ``xt[N, n]`` to the
reference's permuted frequency layout ``xf[n, 2N]`` and back
(N_LAS_BlockSystemHelpers.C:512-609), and ``load`` is one HB load of a group
that holds every sample's instances (N_LOA_HBLoader.C:540-775).
``HBPrecond`` is the block-Jacobi preconditioner (N_LAS_HBBlockJacobiPrecond.C)
factored in FP32 on the matrix cores and solved with FP64 refinement.
"""
import ctypes as C

import numpy as np

from ._abi import check


def _ptr(t):
    if not (t.is_cuda and t.is_contiguous()):
        raise ValueError("a contiguous CUDA tensor is required")
    return t.data_ptr()


def _f64(t, what):
    import torch
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError(f"{what}: a contiguous float64 CUDA tensor is required")
    return t.data_ptr()


class HBPlan:
    def __init__(self, ctx, n_samples, n_unknowns, period):
        self.ctx = ctx
        self.L = ctx.L
        self.N = int(n_samples)
        self.H = (self.N - 1) // 2
        self.n = int(n_unknowns)
        self.period = float(period)
        self.h = C.c_void_p()
        check(self.L.xa_hb_plan_create(ctx.h, self.N, self.n, self.period, C.byref(self.h)),
              "xa_hb_plan_create", self.L)

    def _shapes(self, xt, xf):
        if xt is not None and tuple(xt.shape) not in ((self.N, self.n), (self.N * self.n,)):
            raise ValueError("time-domain array must be [N, n]")
        if xf is not None and tuple(xf.shape) not in ((self.n, 2 * self.N), (2 * self.N * self.n,)):
            raise ValueError("frequency-domain array must be [n, 2N]")

    def ift(self, xf, xt):
        self._shapes(xt, xf)
        check(self.L.xa_hb_ift(self.h, _f64(xf, "xf"), _f64(xt, "xt")), "xa_hb_ift", self.L)

    def dft(self, xt, xf):
        self._shapes(xt, xf)
        check(self.L.xa_hb_dft(self.h, _f64(xt, "xt"), _f64(xf, "xf")), "xa_hb_dft", self.L)

    def add_jwq(self, ff, qf):
        self._shapes(None, ff)
        self._shapes(None, qf)
        check(self.L.xa_hb_add_jwq(self.h, _f64(ff, "ff"), _f64(qf, "qf")), "xa_hb_add_jwq",
              self.L)

    def load(self, master, ss, xf, vectors, nnz_t, ff=None, qf=None, dfdxdvp_f=None,
             dqdxdvp_f=None):
        """One HB load: vectors.nextSol (the time-domain solution, N*n) is
        filled from xf, the time-domain outputs are zeroed and loaded, and the
        frequency-domain vectors given are formed."""
        self._shapes(None, xf)
        opt = lambda t, w: _f64(t, w) if t is not None else None
        check(self.L.xa_hb_load(self.h, master.h, C.byref(ss), _f64(xf, "xf"), C.byref(vectors),
                                int(nnz_t), opt(ff, "ff"), opt(qf, "qf"),
                                opt(dfdxdvp_f, "dfdxdvp_f"), opt(dqdxdvp_f, "dqdxdvp_f")),
              "xa_hb_load", self.L)

    def close(self):
        if self.h:
            self.L.xa_hb_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HBPrecond:
    """Block-Jacobi preconditioner on the CSR pattern (row_ptr, col_idx) of one
    sample's Jacobian."""

    def __init__(self, plan, row_ptr, col_idx):
        self.plan = plan
        self.L = plan.L
        self._rp = np.ascontiguousarray(row_ptr, np.int32)
        self._ci = np.ascontiguousarray(col_idx, np.int32)
        if len(self._rp) != plan.n + 1:
            raise ValueError("row_ptr must have n + 1 entries")
        self.nnz = int(self._rp[-1])
        self.h = C.c_void_p()
        check(self.L.xa_hb_precond_create(plan.h, self.nnz, self._rp.ctypes.data_as(C.c_void_p),
                                          self._ci.ctypes.data_as(C.c_void_p), C.byref(self.h)),
              "xa_hb_precond_create", self.L)

    def factor(self, dfdx_t, dqdx_t):
        """Average the per-sample values ([N * nnz] each), form and factor the
        H + 1 blocks; returns the factors' info codes."""
        info = np.zeros(self.plan.H + 1, np.int32)
        check(self.L.xa_hb_precond_factor(self.h, _f64(dfdx_t, "dFdx_t"), _f64(dqdx_t, "dQdx_t"),
                                          info.ctypes.data_as(C.c_void_p)),
              "xa_hb_precond_factor", self.L)
        return info

    def solve(self, bf, xf, max_refine=8, tol=1e-14):
        """xf = K^-1 bf for every harmonic; returns (rel_fp32, rel_final,
        iterations) per frequency position."""
        N = self.plan.N
        r32 = np.zeros(N)
        rf = np.zeros(N)
        it = np.zeros(N, np.int32)
        check(self.L.xa_hb_precond_solve(self.h, _f64(bf, "bf"), _f64(xf, "xf"), int(max_refine),
                                         float(tol), r32.ctypes.data_as(C.c_void_p),
                                         rf.ctypes.data_as(C.c_void_p),
                                         it.ctypes.data_as(C.c_void_p)),
              "xa_hb_precond_solve", self.L)
        return r32, rf, it

    def close(self):
        if self.h:
            self.L.xa_hb_precond_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lu_f32_factor(ctx, a, m):
    """In-place FP32 LU of a [batch, ld, ld] float32 CUDA tensor (rows and
    columns [m, ld) the identity); returns (ipiv [batch, ld] int32, info)."""
    import torch
    if not (a.is_cuda and a.dtype == torch.float32 and a.is_contiguous() and a.dim() == 3):
        raise ValueError("a must be a contiguous [batch, ld, ld] float32 CUDA tensor")
    batch, ld, ld2 = a.shape
    if ld != ld2:
        raise ValueError("square blocks expected")
    ipiv = torch.zeros((batch, ld), dtype=torch.int32, device=a.device)
    info = torch.zeros(batch, dtype=torch.int32, device=a.device)
    check(ctx.L.xa_lu_f32_factor(ctx.h, batch, int(m), ld, _ptr(a), _ptr(ipiv), _ptr(info)),
          "xa_lu_f32_factor", ctx.L)
    return ipiv, info


def lu_f32_solve(ctx, lu, ipiv, m, b):
    """b [batch, nrhs, ld] float32 overwritten by the solutions."""
    import torch
    if not (b.is_cuda and b.dtype == torch.float32 and b.is_contiguous() and b.dim() == 3):
        raise ValueError("b must be a contiguous [batch, nrhs, ld] float32 CUDA tensor")
    batch, nrhs, ld = b.shape
    check(ctx.L.xa_lu_f32_solve(ctx.h, batch, nrhs, int(m), ld, _ptr(lu), _ptr(ipiv), _ptr(b)),
          "xa_lu_f32_solve", ctx.L)
