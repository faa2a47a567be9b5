"""Fixed-step transient of a netlist circuit on one GPU, with a sparse
Newton solve: every iteration loads all of the circuit's device groups
(``netlist_reader.Circuit``) and solves J dx = rhs by GMRES preconditioned
with ILU(0), on the device.

The Newton residual and Jacobian are formed the way the reference's
integrators form them (backward Euler N_TIA_BackwardDifferentiation15.C:
215-265; the one-step integrator's trapezoid rule N_TIA_OneStep.C:272-335,
463-500, order 1 for the first two steps as N_TIA_OneStep.C:2130,2211
raise it):

* DC operating point: rhs = -F + dFdx*dVp, J = dF/dx (junction
  initialisation on the first iteration, and tranopFlag set, as the
  reference's operating point of a transient, N_ANP_AnalysisManager.C:
  2749-2753, so the loads compute the charges the first step starts from);
* backward Euler: rhs = -((Q - q_n)/h + F) + dFdx*dVp + dQdx*dVp/h,
  J = dF/dx + dQ/dx / h;
* trapezoid: rhs = -((Q - q_n)/h + (F + F_n)/2) + dFdx*dVp/2 + dQdx*dVp/h,
  J = dF/dx / 2 + dQ/dx / h.

F includes the independent sources at the step's time (the sources are
devices of the circuit).  Each voltage source's branch row and its node's
row are exchanged in the solved system, so the incomplete factorisation
finds a pivot on every diagonal.  Newton stops when the largest update of a node
voltage is below vtol volts.  J lives in the circuit's CSR pattern plus
every diagonal (a voltage source's branch row has none of its own); the
reference solves these systems with KLU or preconditioned GMRES
(N_LAS_AztecOOSolver.C)."""
import time

import numpy as np

from .krylov import CSRMatrix, ILU0, fgmres


class NewtonFailure(RuntimeError):
    pass


class SparseTransient:
    def __init__(self, circuit, n_voltages=None, vtol=1e-9, krylov_tol=1e-10, restart=60,
                 max_krylov=600, order="rcm", verbose=False):
        """n_voltages: the unknowns [0, n_voltages) that are node voltages
        (the convergence test reads them; default: the netlist's nodes).
        order: "rcm" solves the system with its unknowns (and rows alike)
        in reverse Cuthill-McKee order of the Jacobian's graph, which keeps
        the incomplete factorisation's fill local; "natural" keeps the LID
        order."""
        import torch
        from . import default_solver_state
        self.torch = torch
        self._ss = default_solver_state
        self.c = circuit
        self.ctx = circuit.ctx
        n = circuit.n_unknowns
        self.n = n
        self.nv = circuit.netlist.n_nodes if n_voltages is None else int(n_voltages)
        self.vtol, self.krylov_tol = vtol, krylov_tol
        self.restart, self.max_krylov = restart, max_krylov
        # rows of the system: a voltage source's branch row and its node's
        # row trade places, so that neither diagonal is structurally zero
        # (the node's KCL row has the branch current at the branch column,
        # the branch row the node voltage at the node column)
        src = np.arange(n, dtype=np.int64)
        for k in range(circuit.n_groups):
            if circuit.group(k)[1] != "vsource":
                continue
            for pos, neg, bra in circuit.group_lids(k).reshape(-1, 3):
                node = pos if pos >= 0 else neg
                if node < 0:
                    continue
                if src[node] != node or src[bra] != bra:
                    raise ValueError("two voltage sources on one node: rows cannot be paired")
                src[node], src[bra] = bra, node
        rp, ci = circuit.csr()
        rows = src[np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))]   # src is its own inverse
        cols = ci.astype(np.int64)
        # symmetric ordering q: position p of the solved system holds unknown
        # q[p] (and the row paired with it)
        if order == "rcm":
            import scipy.sparse as sps
            from scipy.sparse.csgraph import reverse_cuthill_mckee
            g = sps.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
            q = np.asarray(reverse_cuthill_mckee((g + g.T).tocsr(), symmetric_mode=True),
                           np.int64)
        elif order == "natural":
            q = np.arange(n, dtype=np.int64)
        else:
            raise ValueError("order must be 'rcm' or 'natural'")
        qinv = np.empty(n, np.int64)
        qinv[q] = np.arange(n)
        self.verbose = verbose
        self.row_of = self.torch.from_numpy(src[q]).cuda()     # system row p <- circuit row
        self.q = self.torch.from_numpy(q).cuda()
        rows, cols = qinv[rows], qinv[cols]
        ckeys = rows * n + cols
        keys = np.unique(np.concatenate([ckeys, np.arange(n, dtype=np.int64) * (n + 1)]))
        jr = keys // n
        jrp = np.zeros(n + 1, np.int64)
        np.add.at(jrp, jr + 1, 1)
        self.J = CSRMatrix(self.ctx, np.cumsum(jrp).astype(np.int32), (keys % n).astype(np.int32))
        self.map = torch.from_numpy(np.searchsorted(keys, ckeys)).cuda()
        self.ilu = ILU0(self.J)
        f64 = dict(dtype=torch.float64, device="cuda")
        ns, nt, nz = circuit.n_sta, circuit.n_sto, circuit.nnz
        sizes = dict(nextSol=n, currSol=n, f=n, q=n, b=n, dFdxdVp=n, dQdxdVp=n,
                     nextSta=ns, currSta=ns, lastSta=ns, nextStaDeriv=ns,
                     nextSto=nt, currSto=nt, storeLeadF=nt, storeLeadQ=nt, dFdx=nz, dQdx=nz)
        self.v = {k: torch.zeros(max(m, 1), **f64) for k, m in sizes.items()}
        self.outs = ("f", "q", "b", "dFdxdVp", "dQdxdVp", "dFdx", "dQdx", "storeLeadF",
                     "storeLeadQ")
        self.stats = {"loads": 0, "newton": 0, "krylov": 0, "load_s": 0.0, "factor_s": 0.0,
                      "krylov_s": 0.0}

    def load(self, x, flags):
        torch = self.torch
        v = self.v
        v["nextSol"][:self.n].copy_(x)
        for k in self.outs:
            v[k].zero_()
        ss = self._ss(gainScale=1.0, nltermScale=1.0, **flags)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.c.load(ss, **v)
        torch.cuda.synchronize()
        self.stats["load_s"] += time.perf_counter() - t0
        self.stats["loads"] += 1

    def accept(self):
        self.v["currSta"].copy_(self.v["nextSta"])
        self.v["currSto"].copy_(self.v["nextSto"])

    def _newton(self, x, flags_of, system, max_iter):
        torch = self.torch
        v = self.v
        for it in range(max_iter):
            self.load(x, flags_of(it))
            rhs, a, b = system()
            rhs = rhs[self.row_of]
            t0 = time.perf_counter()
            Jv = self.J.values
            Jv.zero_()
            Jv[self.map] = a * v["dFdx"][:len(self.map)] + b * v["dQdx"][:len(self.map)]
            zp = self.ilu.factor()
            if zp >= 0:
                raise NewtonFailure(f"ILU(0): zero pivot in row {zp}")
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            dy, rel, its = fgmres(self.J.matvec, self.ilu.solve, rhs, self.krylov_tol,
                                  self.restart, self.max_krylov)
            dx = torch.empty_like(dy)
            dx[self.q] = dy
            torch.cuda.synchronize()
            self.stats["factor_s"] += t1 - t0
            self.stats["krylov_s"] += time.perf_counter() - t1
            self.stats["krylov"] += its
            self.stats["newton"] += 1
            x_load = x
            x = x + dx
            dmax = float(torch.max(torch.abs(dx[:self.nv])))
            if self.verbose:
                print(f"    Newton {it}: |dx| {dmax:.3e}, GMRES {its} (rel {rel:.1e})", flush=True)
            if rel > 1e3 * self.krylov_tol:
                raise NewtonFailure(f"GMRES stalled at relative residual {rel:.2e}")
            if not np.isfinite(dmax):
                raise NewtonFailure("Newton update is not finite")
            if it > 0 and dmax <= self.vtol:
                return x, x_load, it + 1
        raise NewtonFailure(f"Newton did not converge in {max_iter} iterations")

    def dcop(self, x0, max_iter=100, times=None):
        """DC operating point from x0 (a device tensor of n unknowns)."""
        v = self.v
        extra = dict(times or {})

        def flags(it):
            return dict(newtonIter=it, dcopFlag=1, tranopFlag=1,
                        initJctFlag=1 if it == 0 else 0, currTime=0.0, **extra)

        def system():
            n = self.n
            return -v["f"][:n] + v["dFdxdVp"][:n], 1.0, 0.0

        x, _, its = self._newton(x0, flags, system, max_iter)
        self.accept()
        return x, its

    def run(self, x0, t_stop, h, method="trap", record=None, max_iter=50, progress=None):
        """DC operating point, then fixed steps of h to t_stop.  record:
        unknown indices whose values are kept at every step.  Returns
        (times, recorded[steps+1, len(record)], per-step Newton counts)."""
        if method not in ("be", "trap"):
            raise ValueError("method must be 'be' or 'trap'")
        torch = self.torch
        v = self.v
        n = self.n
        times = dict(startingTimeStep=h, finalTime=t_stop, bpTol=1e-18)
        rec = torch.as_tensor(np.asarray(record if record is not None else [], np.int64)).cuda()
        x, its = self.dcop(x0, times=times)
        newton = [its]
        q_n = v["q"][:n].clone()
        f_n = v["f"][:n].clone()
        out_t = [0.0]
        out_v = [x[rec].cpu().numpy()]
        steps = int(round(t_stop / h))
        for k in range(steps):
            first = k == 0
            tk = (k + 1) * h

            def flags(it, first=first, tk=tk):
                return dict(newtonIter=it, transientFlag=1, pdt=1.0 / h, currTimeStep=h,
                            lastTimeStep=h, currTime=tk, initTranFlag=1 if first else 0,
                            beginIntegrationFlag=1 if first else 0, **times)

            trap = method == "trap" and k >= 2

            def system(q_n=q_n, f_n=f_n, trap=trap):
                f, q = v["f"][:n], v["q"][:n]
                if trap:
                    rhs = (-((q - q_n) / h + 0.5 * f + 0.5 * f_n) + 0.5 * v["dFdxdVp"][:n]
                           + v["dQdxdVp"][:n] / h)
                    return rhs, 0.5, 1.0 / h
                rhs = -((q - q_n) / h + f) + v["dFdxdVp"][:n] + v["dQdxdVp"][:n] / h
                return rhs, 1.0, 1.0 / h

            x, _, its = self._newton(x, flags, system, max_iter)
            self.accept()
            newton.append(its)
            q_n = v["q"][:n].clone()
            f_n = v["f"][:n].clone()
            out_t.append(tk)
            out_v.append(x[rec].cpu().numpy())
            if progress and (k + 1) % progress == 0:
                print(f"  step {k + 1}/{steps}: t = {tk:.3e} s, Newton {its}, "
                      f"GMRES total {self.stats['krylov']}", flush=True)
        return np.asarray(out_t), np.asarray(out_v), np.asarray(newton)
