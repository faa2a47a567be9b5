"""Krylov solves on the device for the circuit-level Newton systems
(harmonic balance, large transients): torch tensors for the vectors, the
operator and the preconditioner supplied by the caller (products through
xa_csr_spmv, the HB transforms, the block-Jacobi or ILU(0) preconditioners of
the library).  The reference solves these systems with AztecOO/Belos GMRES
(N_LAS_AztecOOSolver.C, N_LAS_BelosSolver.C)."""
import ctypes as C

import numpy as np

from ._abi import check


def fgmres(matvec, precond, rhs, tol, restart=40, max_iter=400):
    """Flexible GMRES with right preconditioning (the preconditioned
    directions are kept, so a preconditioner that is not a fixed linear map,
    such as a refined FP32 solve, is allowed): solves A x = rhs to
    |rhs - A x| <= tol |rhs|.  Classical Gram-Schmidt applied twice, Givens
    rotations on the host.  Returns (x, relative residual, iterations)."""
    import torch
    shape = rhs.shape
    b = rhs.reshape(-1)
    bnorm = float(torch.linalg.vector_norm(b))
    x = torch.zeros_like(b)
    if bnorm == 0.0:
        return x.reshape(shape), 0.0, 0
    total = 0
    r = b.clone()
    while total < max_iter:
        beta = float(torch.linalg.vector_norm(r))
        if beta / bnorm <= tol:
            break
        m = min(restart, max_iter - total)
        V = torch.zeros((m + 1, b.numel()), dtype=b.dtype, device=b.device)
        Z = torch.zeros((m, b.numel()), dtype=b.dtype, device=b.device)
        Hm = np.zeros((m + 1, m))
        g = np.zeros(m + 1)
        g[0] = beta
        cs, sn = np.zeros(m), np.zeros(m)
        V[0] = r / beta
        k = 0
        for j in range(m):
            Z[j] = precond(V[j].reshape(shape)).reshape(-1)
            w = matvec(Z[j].reshape(shape)).reshape(-1)
            h = V[:j + 1] @ w
            w = w - V[:j + 1].T @ h
            h2 = V[:j + 1] @ w
            w = w - V[:j + 1].T @ h2
            Hm[:j + 1, j] = (h + h2).cpu().numpy()
            hn = float(torch.linalg.vector_norm(w))
            Hm[j + 1, j] = hn
            if hn > 0:
                V[j + 1] = w / hn
            for i in range(j):
                t = cs[i] * Hm[i, j] + sn[i] * Hm[i + 1, j]
                Hm[i + 1, j] = -sn[i] * Hm[i, j] + cs[i] * Hm[i + 1, j]
                Hm[i, j] = t
            d = np.hypot(Hm[j, j], Hm[j + 1, j])
            cs[j], sn[j] = (1.0, 0.0) if d == 0 else (Hm[j, j] / d, Hm[j + 1, j] / d)
            Hm[j, j] = d
            Hm[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            k = j + 1
            total += 1
            if abs(g[j + 1]) / bnorm <= tol or hn == 0:
                break
        yk = np.linalg.solve(np.triu(Hm[:k, :k]), g[:k])
        x = x + Z[:k].T @ torch.from_numpy(yk).to(b.device)
        r = b - matvec(x.reshape(shape)).reshape(-1)
    rel = float(torch.linalg.vector_norm(r)) / bnorm
    return x.reshape(shape), rel, total


class CSRMatrix:
    """A CSR matrix on the device for xa_csr_spmv (pattern fixed, values a
    float64 tensor that the caller refills, then calls sync()).  Rows longer
    than long_row (a supply's KCL row) are split off into their own CSR and
    multiplied by xa_csr_spmv_long, one block per row, so that one wavefront
    does not walk them alone."""

    def __init__(self, ctx, row_ptr, col_idx, long_row=4096):
        import torch
        self.ctx = ctx
        self.L = ctx.L
        self.n = len(row_ptr) - 1
        self.nnz = int(row_ptr[-1])
        self.row_ptr = np.ascontiguousarray(row_ptr, np.int32)
        self.col_idx = np.ascontiguousarray(col_idx, np.int32)
        self.values = torch.zeros(max(self.nnz, 1), dtype=torch.float64, device="cuda")
        lens = np.diff(self.row_ptr.astype(np.int64))
        longr = np.nonzero(lens > long_row)[0]
        self.split = len(longr) > 0
        if not self.split:
            self.rp = torch.from_numpy(self.row_ptr).cuda()
            self.ci = torch.from_numpy(self.col_idx).cuda()
            return
        rows = np.repeat(np.arange(self.n), lens)
        in_long = np.isin(rows, longr)
        srp = np.zeros(self.n + 1, np.int64)
        np.add.at(srp, rows[~in_long] + 1, 1)
        lrp = np.concatenate([[0], np.cumsum(lens[longr])])
        self.rp = torch.from_numpy(np.cumsum(srp).astype(np.int32)).cuda()
        self.ci = torch.from_numpy(self.col_idx[~in_long]).cuda()
        self.s_src = torch.from_numpy(np.nonzero(~in_long)[0]).cuda()
        self.l_rp = torch.from_numpy(lrp.astype(np.int32)).cuda()
        self.l_ci = torch.from_numpy(self.col_idx[in_long]).cuda()
        self.l_src = torch.from_numpy(np.nonzero(in_long)[0]).cuda()
        self.l_rows = torch.from_numpy(longr).cuda()
        self.s_vals = torch.zeros(max(len(self.s_src), 1), dtype=torch.float64, device="cuda")
        self.l_vals = torch.zeros(max(len(self.l_src), 1), dtype=torch.float64, device="cuda")
        self.l_y = torch.zeros(len(longr), dtype=torch.float64, device="cuda")

    def sync(self):
        """Take the current values into the split parts (after a refill)."""
        if self.split:
            self.s_vals[:len(self.s_src)] = self.values[self.s_src]
            self.l_vals[:len(self.l_src)] = self.values[self.l_src]

    def matvec(self, x, y=None, values=None):
        import torch
        y = torch.empty(self.n, dtype=x.dtype, device=x.device) if y is None else y
        if not self.split:
            vals = self.values if values is None else values
            check(self.L.xa_csr_spmv(self.ctx.h, self.n, self.rp.data_ptr(), self.ci.data_ptr(),
                                     vals.data_ptr(), x.data_ptr(), y.data_ptr()),
                  "xa_csr_spmv", self.L)
            return y
        if values is not None:
            raise ValueError("a split matrix multiplies its synced values")
        check(self.L.xa_csr_spmv(self.ctx.h, self.n, self.rp.data_ptr(), self.ci.data_ptr(),
                                 self.s_vals.data_ptr(), x.data_ptr(), y.data_ptr()),
              "xa_csr_spmv", self.L)
        check(self.L.xa_csr_spmv_long(self.ctx.h, len(self.l_rows), self.l_rp.data_ptr(),
                                      self.l_ci.data_ptr(), self.l_vals.data_ptr(), x.data_ptr(),
                                      self.l_y.data_ptr()), "xa_csr_spmv_long", self.L)
        y[self.l_rows] = self.l_y
        return y


class ILU0:
    """ILU(0) of a CSRMatrix's pattern on the device (xa_ilu0_*)."""

    def __init__(self, mat):
        self.mat = mat
        self.L = mat.L
        self.h = C.c_void_p()
        check(self.L.xa_ilu0_create(mat.ctx.h, mat.n, mat.nnz,
                                    mat.row_ptr.ctypes.data_as(C.c_void_p),
                                    mat.col_idx.ctypes.data_as(C.c_void_p), C.byref(self.h)),
              "xa_ilu0_create", self.L)

    def factor(self, values=None):
        """Factor the matrix's values (or the given ones); returns the row of
        the first zero pivot, or -1."""
        vals = self.mat.values if values is None else values
        zp = C.c_int(-1)
        check(self.L.xa_ilu0_factor(self.h, vals.data_ptr(), C.byref(zp)), "xa_ilu0_factor",
              self.L)
        return zp.value

    def config(self, mode=-1):
        """Set the triangular-solve method (1 level-scheduled kernels, the
        default; 2 the same replayed as a HIP graph; 0 rocSPARSE csrsv; -1
        keep); returns (L levels, U levels)."""
        lv, uv = C.c_int(), C.c_int()
        check(self.L.xa_ilu0_config(self.h, int(mode), C.byref(lv), C.byref(uv)),
              "xa_ilu0_config", self.L)
        return lv.value, uv.value

    def solve(self, b, x=None):
        import torch
        x = torch.empty_like(b) if x is None else x
        check(self.L.xa_ilu0_solve(self.h, b.data_ptr(), x.data_ptr()), "xa_ilu0_solve", self.L)
        return x

    def close(self):
        if self.h:
            self.L.xa_ilu0_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
