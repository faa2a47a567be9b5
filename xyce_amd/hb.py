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
import sys

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


def sample_steps(sources, n_samples, period):
    """`.STEP` lines that give every HB time sample its own copy of the
    circuit with the sources at that sample's values: sources maps a voltage
    source's name to a function of time; sample s sits at t = s * period / N
    (the HB loader sets each sample's time before its load,
    N_LOA_HBLoader.C:618-660).  The netlist gives those sources as DC."""
    t = np.arange(n_samples) * (float(period) / n_samples)
    lines = []
    for name, fn in sources.items():
        vals = np.broadcast_to(np.asarray(fn(t), np.float64), (n_samples,))
        lines.append(f".STEP {name}:DCV0 LIST " + " ".join(repr(float(v)) for v in vals))
    return "\n".join(lines) + "\n"


class HBSolver:
    """Harmonic-balance Newton-Krylov solve on the device.

    The circuit is a netlist built batched over its time samples
    (``netlist_reader.Circuit(ctx, netlist, step=-2)`` with one `.STEP` copy
    per sample, see :func:`sample_steps`): copy s holds sample s's unknowns
    at s*n .. s*n+n-1 and its devices, so one load evaluates every sample.
    The solve follows the reference's HB Newton with a matrix-free Krylov
    solve (N_LOA_HBLoader.C:350-456 for the product, :540-775 for the
    residual, N_LAS_HBBlockJacobiPrecond.C for the preconditioner):

    * residual R(X) = DFT(F(IFT X)) + j w k DFT(Q(IFT X)) in the permuted
      frequency layout (the sources' values enter F);
    * J V = DFT(dF/dx IFT V) + j w k DFT(dQ/dx IFT V), the per-sample
      Jacobians applied by ``xa_csr_spmv`` to the whole block-diagonal
      time-domain system;
    * flexible GMRES (the preconditioner's refined FP32 solve is not a fixed
      linear map) with the block-Jacobi preconditioner on the right;
    * unknowns are the independent half of the spectrum: harmonic 0's real
      part and harmonics 1..H (harmonic -k is the conjugate of k; the
      inverse transform reads only k >= 0).

    The circuit must be built with the voltage limiter off: HB has no
    previous time point for the limiter to refer to."""

    def __init__(self, circuit, period, flags=None):
        import torch
        from . import default_solver_state
        self.torch = torch
        self.c = circuit
        self.ctx = circuit.ctx
        self.L = self.ctx.L
        N = circuit.n_copies
        U = circuit.n_unknowns // N
        if N < 1 or N % 2 == 0 or U * N != circuit.n_unknowns:
            raise ValueError("HB needs an odd number of sample copies (2H + 1)")
        self.N, self.n, self.H = N, U, (N - 1) // 2
        self.period = float(period)
        rp, ci = circuit.csr()
        nnz1 = int(rp[U])
        for s in range(N):
            blk_rp = rp[s * U:(s + 1) * U + 1] - rp[s * U]
            if not (np.array_equal(blk_rp, rp[:U + 1]) and
                    np.array_equal(ci[rp[s * U]:rp[(s + 1) * U]] - s * U, ci[:nnz1])):
                raise ValueError("the sample copies must share one block-diagonal CSR pattern")
        self.plan = HBPlan(self.ctx, N, U, self.period)
        self.pc = HBPrecond(self.plan, rp[:U + 1], ci[:nnz1])
        dev = "cuda"
        self.rp = torch.from_numpy(np.ascontiguousarray(rp, np.int32)).to(dev)
        self.ci = torch.from_numpy(np.ascontiguousarray(ci, np.int32)).to(dev)
        f64 = dict(dtype=torch.float64, device=dev)
        n, ns, nt, nz = circuit.n_unknowns, circuit.n_sta, circuit.n_sto, circuit.nnz
        sizes = dict(nextSol=n, currSol=n, f=n, q=n, b=n, dFdxdVp=n, dQdxdVp=n,
                     nextSta=ns, currSta=ns, lastSta=ns, nextStaDeriv=ns,
                     nextSto=nt, currSto=nt, storeLeadF=nt, storeLeadQ=nt, dFdx=nz, dQdx=nz)
        self.v = {k: torch.zeros(max(s_, 1), **f64) for k, s_ in sizes.items()}
        flags = dict(flags or {})
        flags.setdefault("transientFlag", 1)
        flags.setdefault("newtonIter", 1)
        flags.setdefault("pdt", 1.0 / (self.period / N))
        self.ss = default_solver_state(gainScale=1.0, nltermScale=1.0, **flags)
        self.ff = torch.zeros((U, 2 * N), **f64)
        self.qf = torch.zeros((U, 2 * N), **f64)
        self.tmp_t = [torch.zeros(n, **f64) for _ in range(3)]
        self.tmp_f = torch.zeros((U, 2 * N), **f64)
        self.loads = 0
        self.matvecs = 0
        self.precond_solves = 0

    # ---- layouts ----
    def pack(self, xf):
        """[n, 2N] permuted spectrum -> [n, N] independent part (Re_0, then
        Re_k, Im_k for k = 1..H)."""
        return self.torch.cat([xf[:, 0:1], xf[:, 2:self.N + 1]], dim=1).contiguous()

    def unpack(self, y):
        N, H = self.N, self.H
        xf = self.torch.zeros((self.n, 2 * N), dtype=y.dtype, device=y.device)
        xf[:, 0] = y[:, 0]
        xf[:, 2:N + 1] = y[:, 1:]
        k = self.torch.arange(1, H + 1, device=y.device)
        v = xf.view(self.n, N, 2)
        v[:, N - k, 0] = y[:, 2 * k - 1]
        v[:, N - k, 1] = -y[:, 2 * k]
        return xf

    # ---- HB load and Jacobian product ----
    def residual(self, xf):
        """R(xf) = DFT F + j w k DFT Q; leaves the per-sample dF/dx and dQ/dx
        in the circuit's CSR values and the time-domain solution in nextSol."""
        from . import make_vectors
        v = self.v
        self.plan.ift(xf, v["nextSol"])
        v["currSol"].copy_(v["nextSol"])
        for k in ("f", "q", "b", "dFdxdVp", "dQdxdVp", "dFdx", "dQdx", "storeLeadF",
                  "storeLeadQ"):
            v[k].zero_()
        self.c.load(self.ss, **v)
        self.loads += 1
        self.plan.dft(v["f"], self.ff)
        self.plan.dft(v["q"], self.qf)
        self.plan.add_jwq(self.ff, self.qf)
        return self.ff.clone()

    def _spmv(self, vals, x, y):
        check(self.L.xa_csr_spmv(self.ctx.h, self.c.n_unknowns, self.rp.data_ptr(),
                                 self.ci.data_ptr(), vals.data_ptr(), x.data_ptr(),
                                 y.data_ptr()), "xa_csr_spmv", self.L)

    def matvec(self, y):
        """J applied to a packed direction (the last residual's Jacobians)."""
        vt, gt, ct = self.tmp_t
        self.plan.ift(self.unpack(y), vt)
        self._spmv(self.v["dFdx"], vt, gt)
        self._spmv(self.v["dQdx"], vt, ct)
        out = self.torch.zeros((self.n, 2 * self.N), dtype=y.dtype, device=y.device)
        self.plan.dft(gt, out)
        self.plan.dft(ct, self.tmp_f)
        self.plan.add_jwq(out, self.tmp_f)
        self.matvecs += 1
        return self.pack(out)

    def precond(self, y, refine=2, tol=1e-8):
        bf = self.unpack(y)
        xf = self.torch.zeros_like(bf)
        self.pc.solve(bf, xf, max_refine=refine, tol=tol)
        self.precond_solves += 1
        return self.pack(xf)

    def fgmres(self, rhs, tol, restart=40, max_iter=400):
        from .krylov import fgmres
        return fgmres(self.matvec, self.precond, rhs, tol, restart, max_iter)

    def solve(self, xf0, vtol=1e-9, max_newton=40, krylov_tol=1e-9, restart=40,
              max_krylov=400, max_step=0.5, verbose=False, refactor_every=2):
        """Newton from xf0 ([n, 2N] spectrum) until the largest time-domain
        update is below vtol volts; each update is scaled so that no sample
        moves by more than max_step.  refactor_every: the preconditioner is
        refactored from the current Jacobians every that many Newton
        iterations (GMRES always applies the current ones).  Returns (xf,
        stats)."""
        torch = self.torch
        xf = xf0.clone()
        hist = []
        upd_t = torch.zeros(self.c.n_unknowns, dtype=torch.float64, device="cuda")
        for it in range(max_newton):
            R = self.residual(xf)
            rnorm = float(torch.max(torch.abs(self.pack(R))))
            if it % max(1, int(refactor_every)) == 0:
                info = self.pc.factor(self.v["dFdx"], self.v["dQdx"])
                if np.any(info != 0):
                    raise RuntimeError(f"HB preconditioner: singular block (info {info.tolist()})")
            dy, rel, its = self.fgmres(-self.pack(R), krylov_tol, restart, max_krylov)
            dxf = self.unpack(dy)
            self.plan.ift(dxf, upd_t)
            step = float(torch.max(torch.abs(upd_t)))
            if not np.isfinite(step):
                raise RuntimeError("HB Newton: the update is not finite")
            scale = min(1.0, max_step / step) if step > 0 else 1.0
            xf += scale * dxf
            hist.append({"iter": it, "residual_max": rnorm, "update_max": step * scale,
                         "krylov_iterations": its, "krylov_rel": rel, "damping": scale})
            if verbose:
                print(f"  HB Newton {it}: |R| {rnorm:.3e} |dx| {step:.3e} gmres {its} "
                      f"(rel {rel:.1e}) damping {scale:.3f}", file=sys.stderr, flush=True)
            if scale == 1.0 and step <= vtol:
                R = self.residual(xf)
                return xf, {"converged": True, "newton": it + 1, "history": hist,
                            "final_residual_max": float(torch.max(torch.abs(self.pack(R)))),
                            "loads": self.loads, "matvecs": self.matvecs,
                            "precond_solves": self.precond_solves}
        raise RuntimeError(f"HB Newton did not converge in {max_newton} iterations: {hist[-3:]}")
